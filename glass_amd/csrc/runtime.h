// runtime.h -- host-side runtime of libglass_hip: per-thread HIP context,
// error reporting, and the per-device tables the kernels read (FFT twiddles,
// spacecraft positions).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "../../include/glass_hip.h"

namespace glass {

void set_error(const char* fmt, ...);

#define GH_CHECK(call)                                                                         \
    do                                                                                         \
    {                                                                                          \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess)                                                                  \
        {                                                                                      \
            ::glass::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, \
                               __LINE__);                                                      \
            return GLASS_HIP_ERUNTIME;                                                         \
        }                                                                                      \
    } while (0)

// Growable device / pinned-host buffer owned by a thread context.
struct DevBuf
{
    void* p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes);
    ~DevBuf();
};
struct HostBuf
{
    void* p = nullptr;
    size_t n = 0;
    int ensure(size_t bytes);
    ~HostBuf();
};

// One per host thread: the reference's drop-ins are called concurrently
// from OpenMP threads (apps/src/ucb_mcmc.c:230), so each thread gets its
// own stream and staging memory.
struct ThreadCtx
{
    int device = 0;
    bool ready = false;
    hipStream_t stream = nullptr;
    DevBuf dbuf;
    HostBuf hbuf;
    DevBuf dpart;  // reduction partials
    DevBuf dscr;   // per-workgroup scratch of the BW 2048/4096 derivative kernels
    ~ThreadCtx();
};

int ctx_get(ThreadCtx** out);

// FFT twiddles tw[k] = exp(-2 pi i k / kTwN), k < kTwN, on the current device.
constexpr int kTwN = 4096;
int twiddles(const double2** out);
// Same for the sizes N = 3*2^a: tw3[k] = exp(-2 pi i k / kTw3N).
constexpr int kTw3N = 12288;
int twiddles3(const double2** out);
// And for the long WDM template series, N = 2^a up to 16384.
constexpr int kTw16N = 16384;
int twiddles16(const double2** out);
// And up to 32768 (wavelet templates with a forced layer count).
constexpr int kTw32N = 32768;
int twiddles32(const double2** out);

// Spacecraft positions of the analytic orbit at the npos sample times
// t0 + T*n/npos (npos a power of two), 9 doubles per sample
// (x1,x2,x3,y1,y2,y3,z1,z2,z3), resident on the current device.
int positions(const glass_hip_orbit& o, double T, double t0, int npos, const double** out);
// Spacecraft positions per series length N = kPosSoaMin..kPosSoaMax (powers
// of two) in structure-of-arrays form: the block of N is 9 rows of N
// samples (x1,x2,x3,y1,...,z3 at t0 + T n / N) at pos_soa_offset(N), so the
// threads of a wave read consecutive samples of one coordinate.
constexpr int kPosSoaMin = 32, kPosSoaMax = 4096;
__host__ __device__ constexpr int pos_soa_offset(int N) { return 9 * (N - kPosSoaMin); }
int positions_soa(const glass_hip_orbit& o, double T, double t0, const double** out);

// Host restatement of analytic_orbits (utils/src/glass_lisa.c:77-124), used
// only to fill the positions table.
void analytic_positions(const glass_hip_orbit& o, double t, double* x9);

}  // namespace glass
