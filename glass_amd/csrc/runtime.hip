// runtime.hip -- per-thread HIP contexts, errors, device-resident tables.
#include "runtime.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace glass {

static thread_local std::string g_err;
static thread_local int g_device = 0;

void set_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

int DevBuf::ensure(size_t bytes)
{
    if (bytes <= n) return GLASS_HIP_OK;
    if (p) GH_CHECK(hipFree(p));
    p = nullptr;
    n = 0;
    size_t want = bytes < 4096 ? 4096 : bytes + bytes / 2;
    GH_CHECK(hipMalloc(&p, want));
    n = want;
    return GLASS_HIP_OK;
}

DevBuf::~DevBuf()
{
    if (p) (void)hipFree(p);
}

int HostBuf::ensure(size_t bytes)
{
    if (bytes <= n) return GLASS_HIP_OK;
    if (p) GH_CHECK(hipHostFree(p));
    p = nullptr;
    n = 0;
    size_t want = bytes < 4096 ? 4096 : bytes + bytes / 2;
    GH_CHECK(hipHostMalloc(&p, want, hipHostMallocDefault));
    n = want;
    return GLASS_HIP_OK;
}

HostBuf::~HostBuf()
{
    if (p) (void)hipHostFree(p);
}

ThreadCtx::~ThreadCtx()
{
    if (stream) (void)hipStreamDestroy(stream);
}

static thread_local ThreadCtx t_ctx;

int ctx_get(ThreadCtx** out)
{
    ThreadCtx& c = t_ctx;
    if (!c.ready || c.device != g_device)
    {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= g_device)
        {
            set_error("no HIP device %d (found %d)", g_device, ndev);
            return GLASS_HIP_ENODEV;
        }
        GH_CHECK(hipSetDevice(g_device));
        if (c.stream) (void)hipStreamDestroy(c.stream);
        GH_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        c.device = g_device;
        c.ready = true;
    }
    else
    {
        GH_CHECK(hipSetDevice(c.device));
    }
    *out = &c;
    return GLASS_HIP_OK;
}

static std::mutex g_tab_mu;
static std::map<int, double2*> g_tw, g_tw3, g_tw16, g_tw32;
typedef std::tuple<int, double, double, double, double, double, int> PosKey;
static std::map<PosKey, double*> g_pos;
static std::map<PosKey, double*> g_pos_soa;

static int twiddle_table(std::map<int, double2*>& cache, int ntw, const double2** out)
{
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = cache.find(dev);
    if (it != cache.end())
    {
        *out = it->second;
        return GLASS_HIP_OK;
    }
    // kiss_fft twiddle formula (external/kissfft/kiss_fft.c:356-363)
    std::vector<double2> h(ntw);
    const double pi = 3.141592653589793238462643383279502884197169399375105820974944;
    for (int k = 0; k < ntw; k++)
    {
        double ph = -2 * pi * k / ntw;
        h[k] = make_double2(cos(ph), sin(ph));
    }
    double2* d = nullptr;
    GH_CHECK(hipMalloc(&d, sizeof(double2) * ntw));
    GH_CHECK(hipMemcpy(d, h.data(), sizeof(double2) * ntw, hipMemcpyHostToDevice));
    cache[dev] = d;
    *out = d;
    return GLASS_HIP_OK;
}

int twiddles(const double2** out) { return twiddle_table(g_tw, kTwN, out); }
int twiddles3(const double2** out) { return twiddle_table(g_tw3, kTw3N, out); }
int twiddles16(const double2** out) { return twiddle_table(g_tw16, kTw16N, out); }
int twiddles32(const double2** out) { return twiddle_table(g_tw32, kTw32N, out); }

void analytic_positions(const glass_hip_orbit& o, double t, double* p)
{
#pragma clang fp contract(off)
    const double PI2 = 6.283185307179586, YEAR = 3.15581497632e7, AU = 1.49597870700e11;
    const double SQ3 = 1.73205080757;
    double alpha = PI2 * t / YEAR;
    double sa = sin(alpha), ca = cos(alpha);
    double sa2 = sa * sa, ca2 = ca * ca, saca = sa * ca;
    double AUca = AU * ca, AUsa = AU * sa;
    // beta = 0, 2pi/3, 4pi/3 with the reference's literal values (glass_lisa.c:100-116)
    static const double SB[3] = {0.0, 0.866025403784439, -0.866025403784438};
    static const double CB[3] = {1.0, -0.5, -0.5};
    for (int i = 0; i < 3; i++)
    {
        double sb = SB[i], cb = CB[i];
        p[i] = AUca + o.R * (saca * sb - (1. + sa2) * cb);
        p[3 + i] = AUsa + o.R * (saca * cb - (1. + ca2) * sb);
        p[6 + i] = -SQ3 * o.R * (ca * cb + sa * sb);
    }
}

int positions(const glass_hip_orbit& o, double T, double t0, int npos, const double** out)
{
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    PosKey key(dev, o.L, o.fstar, o.R, T, t0, npos);
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_pos.find(key);
    if (it != g_pos.end())
    {
        *out = it->second;
        return GLASS_HIP_OK;
    }
    std::vector<double> h((size_t)npos * 9);
    for (int n = 0; n < npos; n++)
    {
        double t;
        {
#pragma clang fp contract(off)
            t = t0 + T * (double)(n) / (double)npos;
        }
        analytic_positions(o, t, &h[(size_t)n * 9]);
    }
    double* d = nullptr;
    GH_CHECK(hipMalloc(&d, sizeof(double) * h.size()));
    GH_CHECK(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    g_pos[key] = d;
    *out = d;
    return GLASS_HIP_OK;
}

int positions_soa(const glass_hip_orbit& o, double T, double t0, const double** out)
{
    int dev = 0;
    GH_CHECK(hipGetDevice(&dev));
    PosKey key(dev, o.L, o.fstar, o.R, T, t0, kPosSoaMax);
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_pos_soa.find(key);
    if (it != g_pos_soa.end())
    {
        *out = it->second;
        return GLASS_HIP_OK;
    }
    // sample n of an N-sample series is sample n (kPosSoaMax / N) of the
    // finest grid: t0 + T n / N, the same double (the ratio is a power of two)
    std::vector<double> fine((size_t)kPosSoaMax * 9);
    for (int n = 0; n < kPosSoaMax; n++)
    {
        double t;
        {
#pragma clang fp contract(off)
            t = t0 + T * (double)(n) / (double)kPosSoaMax;
        }
        analytic_positions(o, t, &fine[(size_t)n * 9]);
    }
    std::vector<double> h((size_t)9 * (2 * kPosSoaMax - kPosSoaMin));
    for (int N = kPosSoaMin; N <= kPosSoaMax; N *= 2)
    {
        double* blk = h.data() + pos_soa_offset(N);
        for (int n = 0; n < N; n++)
            for (int k = 0; k < 9; k++) blk[(size_t)k * N + n] = fine[(size_t)n * (kPosSoaMax / N) * 9 + k];
    }
    double* d = nullptr;
    GH_CHECK(hipMalloc(&d, sizeof(double) * h.size()));
    GH_CHECK(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    g_pos_soa[key] = d;
    *out = d;
    return GLASS_HIP_OK;
}

}  // namespace glass

extern "C" {

void glass_hip_analytic_orbit(glass_hip_orbit* o)
{
    // utils/src/glass_lisa.c:125-136 with LARM, CLIGHT, AU, SQ3
    const double LARM = 2.5e9, CLIGHT = 299792458., AU = 1.49597870700e11, SQ3 = 1.73205080757;
    o->L = LARM;
    o->fstar = CLIGHT / (2.0 * M_PI * LARM);
    double ecc = LARM / (2.0 * SQ3 * AU);
    o->R = AU * ecc;
}

int glass_hip_set_device(int device)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    {
        glass::set_error("invalid device %d (found %d)", device, ndev);
        return GLASS_HIP_ENODEV;
    }
    glass::g_device = device;
    return GLASS_HIP_OK;
}

const char* glass_hip_last_error(void) { return glass::g_err.c_str(); }

}  // extern "C"
