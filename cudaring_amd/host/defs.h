// defs.h -- constants of the DPM2D host API (values of defs.h:13-23 in farceri/cudaRing).
#ifndef DPM_HOST_DEFS_H
#define DPM_HOST_DEFS_H

#include <vector>

const double PI = 3.141592653589793238462643383279502884197;
const long MAXDIM = 2;
const long precision = 14;         // digits written by ioDPMFile
const double WCAcut = 1.122462048;  // 2^(1/6) as used by the WCA potential

// FIRE defaults
const double alpha0 = 0.2;
const double finc = 1.1;
const double fdec = 0.5;
const double falpha = 0.99;

// The reference passes thrust::host_vector<T>& at this boundary; here host data are
// std::vector<T>.  `dpm::host_vector<T>` names the same type for ported driver code.
namespace dpm {
template <class T>
using host_vector = std::vector<T>;
}

#endif
