// common.h -- enumerations shared by the drop-in solver classes.
// Same enumerators, in the same order, as the reference's include/common.h:4-11 so that
// callers compile unchanged (BC2D values are also the ASD2_BC_* codes of asd2_hip.h).
#ifndef ASD2_COMMON_H
#define ASD2_COMMON_H

enum class TIMEORDERENUM { EULER, RK2, RK3 };
enum class BC2D { PERIODIC, NEUMANN, DIRICHLET, AXISYM, WALL, INLET, OUTLET, PRESSURE, LSFREE, CUSTOM };
enum class BC3D { PERIODIC, NEUMANN, DIRICHLET, WALL, INLET, OUTLET, PRESSURE, LSFREE, CUSTOM };
enum class PLTTYPE { ASCII, BINARY, BOTH, NONE };
enum class POISSONTYPE { MKL, CG, GS, BICGSTAB, ICPCG };
enum class GAXISENUM2D { X, Y, Z };
enum class GAXISENUM2DAXISYM { R, Z };
enum class GAXISENUM3D { X, Y, Z };

#endif
