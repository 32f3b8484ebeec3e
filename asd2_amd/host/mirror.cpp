// mirror.cpp -- device mirrors of the callers' std::vector<double> arrays (resident mode of the
// drop-in classes, include/asd2_resident.h).  No reference counterpart: the reference computes on
// the host vectors themselves.
#include <mutex>
#include <unordered_map>

#include "asd2_resident.h"
#include "host_util.h"

namespace asd2host {

namespace {

struct Entry {
    std::unique_ptr<DevVec> d;
    bool host_newer = false;   // the host vector holds changes the device copy lacks
    bool dev_newer = false;    // the device copy holds changes the host vector lacks
    asd2_ctx* writer = nullptr;  // context whose stream last wrote the device copy
};

struct Table {
    std::mutex mu;
    bool resident = false;
    std::unordered_map<const double*, Entry> map;
    asd2::TransferCounts counts{};
};

// never destroyed: device buffers must not be freed by a static destructor after the HIP
// runtime may have shut down
Table& table() {
    static Table* t = new Table;
    return *t;
}

// wait for the last writer when another context touches the mirror
void order_after(Entry& e, asd2_ctx* ctx) {
    if (e.writer && e.writer != ctx) {
        check(asd2_synchronize(e.writer), "mirror: synchronize");
        e.writer = nullptr;
    }
}

void download(Entry& e, const std::vector<double>& h) {
    if (e.writer) {
        check(asd2_synchronize(e.writer), "mirror: synchronize");
        e.writer = nullptr;
    }
    // the mirror is the vector's data seen from the device: writing it back into the caller's
    // buffer does not change the vector object (size and capacity stay)
    double* dst = const_cast<double*>(h.data());
    check(asd2_memcpy_d2h(dst, e.d->p(), h.size() * 8), "mirror: download");
    count_transfer(false, h.size() * 8);
    e.dev_newer = false;
}

}  // namespace

// called with the table's lock held (from the mirror functions) or without (the classes' scratch
// buffers); the classes are single-threaded, as the reference's are
void count_transfer(bool upload, size_t bytes) {
    Table& t = table();
    if (upload) {
        t.counts.uploads++;
        t.counts.upload_bytes += static_cast<long long>(bytes);
    } else {
        t.counts.downloads++;
        t.counts.download_bytes += static_cast<long long>(bytes);
    }
}

namespace mirror {

bool resident() { return table().resident; }

double* find(asd2_ctx* ctx, const std::vector<double>& h) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.map.find(h.data());
    if (it == t.map.end() || it->second.d->size() != h.size()) return nullptr;
    Entry& e = it->second;
    order_after(e, ctx);
    if (e.host_newer) {
        e.d->put(h);
        e.host_newer = false;
    }
    return e.d->p();
}

double* acquire(asd2_ctx* ctx, const std::vector<double>& h, bool current) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.map.find(h.data());
    if (it != t.map.end() && it->second.d->size() != h.size()) {
        t.map.erase(it);  // the buffer was reused for a vector of another size
        it = t.map.end();
    }
    if (it == t.map.end()) {
        Entry e;
        e.d.reset(new DevVec(h.size()));
        e.host_newer = true;
        it = t.map.emplace(h.data(), std::move(e)).first;
    }
    Entry& e = it->second;
    order_after(e, ctx);
    if (e.host_newer) {
        if (current) e.d->put(h);
        e.host_newer = false;
    }
    return e.d->p();
}

void written(asd2_ctx* ctx, const std::vector<double>& h) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.map.find(h.data());
    if (it == t.map.end()) return;
    it->second.dev_newer = true;
    it->second.host_newer = false;
    it->second.writer = ctx;
}

void sync_host(const std::vector<double>& h) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.map.find(h.data());
    if (it == t.map.end() || !it->second.dev_newer) return;
    download(it->second, h);
}

void forget(const std::vector<double>& h) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.map.find(h.data());
    if (it == t.map.end()) return;
    if (it->second.writer) check(asd2_synchronize(it->second.writer), "mirror: synchronize");
    t.map.erase(it);
}

void release(asd2_ctx* ctx) {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    bool waited = false;
    for (auto& kv : t.map)
        if (kv.second.writer == ctx) {
            if (!waited) check(asd2_synchronize(ctx), "mirror: synchronize");
            waited = true;
            kv.second.writer = nullptr;
        }
}

}  // namespace mirror
}  // namespace asd2host

namespace asd2 {

using asd2host::check;

void SetDeviceResident(bool on) {
    if (!on) SyncHostAll();
    auto& t = asd2host::table();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!on) {
        for (auto& kv : t.map)
            if (kv.second.writer) check(asd2_synchronize(kv.second.writer), "mirror: synchronize");
        t.map.clear();
    }
    t.resident = on;
}

bool DeviceResident() { return asd2host::mirror::resident(); }

void SyncHost(std::vector<double>& v) { asd2host::mirror::sync_host(v); }

void SyncHostAll() {
    auto& t = asd2host::table();
    std::lock_guard<std::mutex> lk(t.mu);
    // the keys are the vectors' buffers; a mirror's size is its vector's
    for (auto& kv : t.map)
        if (kv.second.dev_newer) {
            auto& e = kv.second;
            if (e.writer) {
                check(asd2_synchronize(e.writer), "mirror: synchronize");
                e.writer = nullptr;
            }
            check(asd2_memcpy_d2h(const_cast<double*>(kv.first), e.d->p(), e.d->size() * 8), "mirror: download");
            asd2host::count_transfer(false, e.d->size() * 8);
            e.dev_newer = false;
        }
}

void ModifiedHost(const std::vector<double>& v) {
    auto& t = asd2host::table();
    std::lock_guard<std::mutex> lk(t.mu);
    auto it = t.map.find(v.data());
    if (it == t.map.end()) return;
    it->second.host_newer = true;
    it->second.dev_newer = false;
}

void Forget(const std::vector<double>& v) { asd2host::mirror::forget(v); }

TransferCounts Transfers() {
    auto& t = asd2host::table();
    std::lock_guard<std::mutex> lk(t.mu);
    TransferCounts c = t.counts;
    c.mirrors = static_cast<long long>(t.map.size());
    return c;
}

}  // namespace asd2
