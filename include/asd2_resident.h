// asd2_resident.h -- device residency of the drop-in classes' std::vector arguments.
//
// No reference counterpart: the reference's classes (include/levelset2d.h, mac2d.h) take host
// std::vector<double>& on every call and compute on them in place.  By default the drop-in
// classes keep that contract literally: every call uploads the vectors it reads and downloads the
// vectors it writes, so the caller's host vectors are always current.
//
// In resident mode the device copy of a vector outlives the call:
//   * a vector a method writes (ls of Solve_LevelSet_2D, m_u / m_v of UpdateVel, uhat / vhat, ps,
//     the in-place overloads' outputs ...) gets a device mirror that holds its newest contents;
//     the host vector is NOT refreshed until SyncHost;
//   * a vector that has a mirror is read from it without an upload;
//   * a vector with no mirror is uploaded on every call that reads it, as in the default mode.
// The caller's part of the contract (as with any host/device dual view):
//   * before reading a vector on the host, call SyncHost(v) (or SyncHostAll());
//   * after changing a mirrored vector on the host, call ModifiedHost(v): the next call uploads it;
//   * before a mirrored vector is destroyed or reallocated (resize beyond capacity, assignment
//     from a vector of another size, move-assignment), call Forget(v): mirrors are keyed by the
//     vector's buffer, and a new vector that reuses the address would otherwise see stale data.
// The classes' own members (MACSolver2D::m_u, m_v, m_ps, m_p, m_kappa, ...) are forgotten by the
// classes themselves when they are destroyed or reallocated.  SetDeviceResident(false) brings
// every host vector up to date and drops all mirrors.
#pragma once

#include <vector>

namespace asd2 {

void SetDeviceResident(bool on);
bool DeviceResident();
// bring v's host copy up to date from its mirror (no-op without a mirror or when current)
void SyncHost(std::vector<double>& v);
void SyncHostAll();
// v was changed on the host: its mirror (if any) is refreshed by the next call that reads it
void ModifiedHost(const std::vector<double>& v);
// drop v's mirror (its device contents are discarded; call SyncHost first to keep them)
void Forget(const std::vector<double>& v);
// number of mirrors and bytes moved between host and device by the classes since the process
// started (uploads, downloads): lets a driver check what a time step transfers
struct TransferCounts {
    long long mirrors;
    long long uploads, upload_bytes;
    long long downloads, download_bytes;
};
TransferCounts Transfers();

}  // namespace asd2
