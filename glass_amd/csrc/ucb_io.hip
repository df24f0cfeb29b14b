// ucb_io.hip -- the reference's UCB file formats (host code): readers of
// the verification-binary and injection files, writers and readers of the
// chain, parameter, likelihood, temperature and checkpoint files, and a
// checkpoint of a batched plan's device state.
//
// Text goes through the C library's printf / scanf with the reference's
// format strings, so the bytes written and the values read are those of
// the reference for the same inputs:
//   print_source_params / scan_source_params   ucb/src/glass_ucb_io.c:598-636
//   print_chain_state / scan_chain_state       utils/src/glass_sampler.c:138-161
//   print_psd_state / scan_psd_state           ucb/src/glass_ucb_io.c:566-589
//   print_chain_files (likelihood, temperature) ucb/src/glass_ucb_io.c:494-509
//   save_chain_state / restore_chain_state     ucb/src/glass_ucb_io.c:418-490
//   map_array_to_ucb_params / map_ucb_params_to_array  ucb/src/glass_ucb_model.c:25-53
//   GetVerificationBinary, UCBInjectVerificationSource ucb/src/glass_ucb_data.c:56-160,627-673
//   UCBInjectSimulatedSource's source file     ucb/src/glass_ucb_data.c:270-384
//   chirpmass, amplitude                       utils/src/glass_math.c:298-329
//   rand_r_U_0_1                               utils/src/glass_utils.c:19-22
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/glass_hip_io.h"
#include "fd_batch_hooks.h"
#include "runtime.h"

namespace {

using glass::set_error;

// utils/src/glass_constants.h:54-66
constexpr double kTSUN = 4.92569043916e-6;
constexpr double kPC = 3.0856775815e16;
constexpr double kCLIGHT = 299792458.;
constexpr int kNP = 8;  // parameters per source (UCB_MODEL_NP)

// A Source's named parameters (map_array_to_ucb_params)
struct Named
{
    double f0, costheta, phi, amp, cosi, psi, phiref, dfdt;
};

Named to_named(const double* p, double T)
{
    return Named{p[0] / T, p[1], p[2], exp(p[3]), p[4], p[5], p[6], p[7] / (T * T)};
}

void to_array(const Named& s, double T, double* p)
{
    p[0] = s.f0 * T;
    p[1] = s.costheta;
    p[2] = s.phi;
    p[3] = log(s.amp);
    p[4] = s.cosi;
    p[5] = s.psi;
    p[6] = s.phiref;
    p[7] = s.dfdt * T * T;
}

// snprintf into buf at *len; false when it does not fit
bool put(char* buf, int cap, int* len, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    const int room = cap - *len;
    const int n = vsnprintf(room > 0 ? buf + *len : nullptr, room > 0 ? (size_t)room : 0, fmt, ap);
    va_end(ap);
    if (n < 0 || n >= room) return false;
    *len += n;
    return true;
}

int too_small(const char* what)
{
    set_error("%s: buffer too small", what);
    return -1;
}

bool source_text(const double* params, double T, char* buf, int cap, int* len)
{
    const Named s = to_named(params, T);
    return put(buf, cap, len, "%.16g ", s.f0) && put(buf, cap, len, "%.12g ", s.dfdt) &&
           put(buf, cap, len, "%.12g ", s.amp) && put(buf, cap, len, "%.12g ", s.phi) &&
           put(buf, cap, len, "%.12g ", s.costheta) && put(buf, cap, len, "%.12g ", s.cosi) &&
           put(buf, cap, len, "%.12g ", s.psi) && put(buf, cap, len, "%.12g ", s.phiref);
}

bool chain_state_text(int step, int nlive, double logL, double logLnorm, double t0, char* buf, int cap, int* len)
{
    return put(buf, cap, len, "%i ", step) && put(buf, cap, len, "%i ", nlive) &&
           put(buf, cap, len, "%lg ", logL) && put(buf, cap, len, "%lg ", logLnorm) &&
           put(buf, cap, len, "%.12g ", t0) && put(buf, cap, len, "\n");
}

bool psd_state_text(int step, double logL, double logLnorm, const double* eta, int neta, char* buf, int cap,
                    int* len)
{
    if (!put(buf, cap, len, "%i ", step) || !put(buf, cap, len, "%lg %lg ", logL, logLnorm)) return false;
    for (int n = 0; n < neta; n++)
        if (!put(buf, cap, len, "%lg ", eta[n])) return false;
    return put(buf, cap, len, "\n");
}

// scan_source_params on a stream, then map_ucb_params_to_array.  Every
// conversion must succeed (the reference fails only when none does, so a
// truncated file would pass it with values never read).
bool scan_source(FILE* f, double T, double* params)
{
    Named s{};
    int check = 0;
    check += fscanf(f, "%lg", &s.f0) == 1;
    check += fscanf(f, "%lg", &s.dfdt) == 1;
    check += fscanf(f, "%lg", &s.amp) == 1;
    check += fscanf(f, "%lg", &s.phi) == 1;
    check += fscanf(f, "%lg", &s.costheta) == 1;
    check += fscanf(f, "%lg", &s.cosi) == 1;
    check += fscanf(f, "%lg", &s.psi) == 1;
    check += fscanf(f, "%lg", &s.phiref) == 1;
    if (check != kNP) return false;
    to_array(s, T, params);
    return true;
}

}  // namespace

extern "C" int glass_hip_format_source_params(const double* params, double T, char* buf, int cap)
{
    if (!params || !buf || cap <= 0 || !(T > 0))
    {
        set_error("format_source_params: bad arguments");
        return -1;
    }
    int len = 0;
    buf[0] = 0;
    return source_text(params, T, buf, cap, &len) ? len : too_small("format_source_params");
}

extern "C" int glass_hip_format_chain_state(int step, int nlive, double logL, double logLnorm, double t0, char* buf,
                                            int cap)
{
    if (!buf || cap <= 0)
    {
        set_error("format_chain_state: bad arguments");
        return -1;
    }
    int len = 0;
    buf[0] = 0;
    return chain_state_text(step, nlive, logL, logLnorm, t0, buf, cap, &len) ? len
                                                                             : too_small("format_chain_state");
}

extern "C" int glass_hip_format_psd_state(int step, double logL, double logLnorm, const double* eta, int neta,
                                          char* buf, int cap)
{
    if (!buf || cap <= 0 || neta < 0 || (neta && !eta))
    {
        set_error("format_psd_state: bad arguments");
        return -1;
    }
    int len = 0;
    buf[0] = 0;
    return psd_state_text(step, logL, logLnorm, eta, neta, buf, cap, &len) ? len : too_small("format_psd_state");
}

extern "C" int glass_hip_format_value_line(int step, const double* values, int n, char* buf, int cap)
{
    if (!buf || cap <= 0 || n < 0 || (n && !values))
    {
        set_error("format_value_line: bad arguments");
        return -1;
    }
    int len = 0;
    buf[0] = 0;
    if (!put(buf, cap, &len, "%i ", step)) return too_small("format_value_line");
    for (int i = 0; i < n; i++)
        if (!put(buf, cap, &len, "%lg ", values[i])) return too_small("format_value_line");
    return put(buf, cap, &len, "\n") ? len : too_small("format_value_line");
}

extern "C" int glass_hip_scan_source_params(const char* text, double T, double* params, int* consumed)
{
    if (!text || !params || !(T > 0))
    {
        set_error("scan_source_params: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    Named s{};
    int n = 0;
    const int got = sscanf(text, "%lg %lg %lg %lg %lg %lg %lg %lg%n", &s.f0, &s.dfdt, &s.amp, &s.phi, &s.costheta,
                           &s.cosi, &s.psi, &s.phiref, &n);
    if (got != kNP)
    {
        set_error("scan_source_params: %d of 8 values read", got < 0 ? 0 : got);
        return GLASS_HIP_EINVAL;
    }
    to_array(s, T, params);
    if (consumed) *consumed = n;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_read_verification_binary(const char* path, double T, int NFFT, unsigned iseed,
                                                  double* params, int* qmin, double* fmin, double* fmax)
{
    if (!path || !params || !(T > 0) || NFFT <= 0)
    {
        set_error("read_verification_binary: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    FILE* f = fopen(path, "r");
    if (!f)
    {
        set_error("read_verification_binary: cannot open %s", path);
        return GLASS_HIP_EINVAL;
    }
    char header[1024];
    double f0, dfdt, costheta, phi, m1, m2, cosi, D;
    const bool ok = fgets(header, sizeof(header), f) != nullptr &&
                    fscanf(f, "%lg %lg %lg %lg %lg %lg %lg %lg", &f0, &dfdt, &costheta, &phi, &m1, &m2, &cosi, &D) == 8;
    fclose(f);
    if (!ok)
    {
        set_error("read_verification_binary: error reading %s", path);
        return GLASS_HIP_EINVAL;
    }
    D *= 1000.0;  // kpc -> pc
    unsigned r = iseed;
    const double phi0 = (double)rand_r(&r) / (double)RAND_MAX * M_PI * 2.;
    const double psi = (double)rand_r(&r) / (double)RAND_MAX * M_PI / 4.;
    const double Mc = pow(m1 * m2, 3. / 5.) / pow(m1 + m2, 1. / 5.);
    const double M = Mc * kTSUN, dL = D * kPC / kCLIGHT;
    const double amp = 2. * pow(pow(M, 5) * pow(M_PI * f0, 2), 1. / 3.) / dL;
    const double lo = f0 - (NFFT / 2) / T;
    const int q = (int)(lo * T);
    if (qmin) *qmin = q;
    if (fmin) *fmin = q / T;
    if (fmax) *fmax = (q + NFFT) / T;
    to_array(Named{f0, costheta, phi, amp, cosi, psi, phi0, dfdt}, T, params);
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_read_injection_file(const char* path, double T, int max, double* params, int* n)
{
    if (!path || max < 0 || (max && !params) || !n || !(T > 0))
    {
        set_error("read_injection_file: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    FILE* f = fopen(path, "r");
    if (!f)
    {
        set_error("read_injection_file: cannot open %s", path);
        return GLASS_HIP_EINVAL;
    }
    double f0, dfdt, theta, phi, amp, iota, psi, phi0;
    // the reference's count: passes of its read loop until end of file,
    // less one (a file ending in a newline counts its lines)
    int N = 0;
    while (!feof(f))
    {
        if (!fscanf(f, "%lg %lg %lg %lg %lg %lg %lg %lg", &f0, &dfdt, &theta, &phi, &amp, &iota, &psi, &phi0))
        {
            fclose(f);
            set_error("read_injection_file: error reading %s", path);
            return GLASS_HIP_EINVAL;
        }
        N++;
    }
    rewind(f);
    N--;
    for (int k = 0; k < N; k++)
    {
        // each counted source must convert all eight fields
        if (fscanf(f, "%lg %lg %lg %lg %lg %lg %lg %lg", &f0, &dfdt, &theta, &phi, &amp, &iota, &psi, &phi0) != 8)
        {
            fclose(f);
            set_error("read_injection_file: error reading %s", path);
            return GLASS_HIP_EINVAL;
        }
        // map the polarization angle into [0, pi], keeping its relation to phi0
        if (psi > M_PI) psi -= M_PI;
        if (phi0 > 2. * M_PI) phi0 -= 2. * M_PI;
        if (k < max)
            to_array(Named{f0, cos(M_PI / 2. - theta), phi, amp, cos(iota), psi, phi0, dfdt}, T,
                     params + (size_t)k * kNP);
    }
    fclose(f);
    *n = N;
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_write_chain_state_file(const char* path, double logLmax, int step, int nlive, double logL,
                                                double logLnorm, double t0, const double* eta, int neta,
                                                const double* params, double T)
{
    if (!path || nlive < 0 || (nlive && !params) || neta < 0 || (neta && !eta) || !(T > 0))
    {
        set_error("write_chain_state_file: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    std::vector<char> buf(256 + 32 * (size_t)neta + 256 * (size_t)nlive);
    const int cap = (int)buf.size();
    int len = 0;
    bool ok = put(buf.data(), cap, &len, "%.12g\n", logLmax) &&
              chain_state_text(step, nlive, logL, logLnorm, t0, buf.data(), cap, &len) &&
              psd_state_text(step, logL, logLnorm, eta, neta, buf.data(), cap, &len);
    for (int i = 0; ok && i < nlive; i++)
        ok = source_text(params + (size_t)i * kNP, T, buf.data(), cap, &len) && put(buf.data(), cap, &len, "\n");
    if (!ok)
    {
        set_error("write_chain_state_file: formatting failed");
        return GLASS_HIP_EINVAL;
    }
    FILE* f = fopen(path, "w");
    if (!f)
    {
        set_error("write_chain_state_file: cannot open %s", path);
        return GLASS_HIP_EINVAL;
    }
    const bool wrote = fwrite(buf.data(), 1, (size_t)len, f) == (size_t)len;
    if (fclose(f) != 0 || !wrote)
    {
        set_error("write_chain_state_file: write to %s failed", path);
        return GLASS_HIP_EINVAL;
    }
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_read_chain_state_file(const char* path, int maxsrc, int neta, double T, double* logLmax,
                                               int* step, int* nlive, double* logL, double* logLnorm, double* t0,
                                               double* eta, double* params)
{
    if (!path || maxsrc < 0 || (maxsrc && !params) || neta < 0 || (neta && !eta) || !logLmax || !step || !nlive ||
        !logL || !logLnorm || !t0 || !(T > 0))
    {
        set_error("read_chain_state_file: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    FILE* f = fopen(path, "r");
    if (!f)
    {
        set_error("read_chain_state_file: cannot open %s", path);
        return GLASS_HIP_EINVAL;
    }
    auto fail = [&](const char* what) {
        fclose(f);
        set_error("read_chain_state_file: error reading the %s of %s", what, path);
        return GLASS_HIP_EINVAL;
    };
    // the reference's reads, each required to convert all its fields
    if (fscanf(f, "%lg", logLmax) != 1) return fail("maximum logL");
    // scan_chain_state
    int check = 0;
    check += fscanf(f, "%i", step) == 1;
    check += fscanf(f, "%i", nlive) == 1;
    check += fscanf(f, "%lg", logL) == 1;
    check += fscanf(f, "%lg", logLnorm) == 1;
    check += fscanf(f, "%lg", t0) == 1;
    if (check != 5) return fail("chain state");
    // scan_psd_state
    check = 0;
    check += fscanf(f, "%i ", step) == 1;
    check += fscanf(f, "%lg %lg ", logL, logLnorm) == 2;
    for (int n = 0; n < neta; n++) check += fscanf(f, "%lg", &eta[n]) == 1;
    if (check != 2 + neta) return fail("noise state");
    if (*nlive < 0 || *nlive > maxsrc)
    {
        fclose(f);
        set_error("read_chain_state_file: %d live sources in %s, room for %d", *nlive, path, maxsrc);
        return GLASS_HIP_EINVAL;
    }
    for (int i = 0; i < *nlive; i++)
        if (!scan_source(f, T, params + (size_t)i * kNP)) return fail("source parameters");
    fclose(f);
    return GLASS_HIP_OK;
}

extern "C" int glass_hip_fd_batch_save_chain_states(glass_hip_fd_batch* plan, const char* dir, int step,
                                                    double logLmax, const double* logLnorm, double t0,
                                                    const double* eta, int neta)
{
    if (!plan || !dir || neta < 0 || (neta && !eta))
    {
        set_error("fd_batch_save_chain_states: bad arguments");
        return GLASS_HIP_EINVAL;
    }
    const glass::FdBatchView v = glass::fd_batch_view(plan);
    GH_CHECK(hipSetDevice(v.device));
    std::vector<int> nsrc(v.C);
    std::vector<double> logl(v.C), sp((size_t)v.C * v.maxsrc * kNP);
    GH_CHECK(hipMemcpy(nsrc.data(), v.c_nsrc, sizeof(int) * v.C, hipMemcpyDeviceToHost));
    if (int rc = glass_hip_fd_batch_state(plan, logl.data(), sp.data())) return rc;
    std::vector<char> path(4096);
    for (int c = 0; c < v.C; c++)
    {
        if (snprintf(path.data(), path.size(), "%s/chain_state_%i.dat", dir, c) >= (int)path.size())
        {
            set_error("fd_batch_save_chain_states: directory name too long");
            return GLASS_HIP_EINVAL;
        }
        if (int rc = glass_hip_write_chain_state_file(path.data(), logLmax, step, nsrc[c], logl[c],
                                                      logLnorm ? logLnorm[c] : 0.0, t0, eta, neta,
                                                      sp.data() + (size_t)c * v.maxsrc * kNP, v.T))
            return rc;
    }
    return GLASS_HIP_OK;
}
