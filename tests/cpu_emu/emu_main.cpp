// Scenarios that drive the library's C ABI, built for the CPU against the
// emulated HIP runtime (hip/hip_runtime.h) under AddressSanitizer and
// UndefinedBehaviorSanitizer, so that every kernel's indexing of device
// memory and of its workgroup arrays is checked on the CPU.  Each scenario
// also compares its result with the oracle (oracle/sc_oracle.c), so a kernel
// that stays in bounds but reads the wrong cells fails too.
//
// Usage: emu_main [scenario-substring]   (default: all)
// Prints one line per scenario and exits 0 when every one passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/sharpclaw_amd.h"

extern "C" {
double orc_ts(const sc_params *p, double *q, const double *qold, double *r1, double *r2, double dt, double told,
              long long *nevals);
int orc_sharpclaw(const sc_params *p, const sc_run_ctl *c, double *q, double tstart, double tend, double *dt_io,
                  sc_run_stats *st);
void orc_set_aux(const double *aux);
void orc_swirl_setaux(const sc_params *p, double *aux);
}

namespace {

struct Rng {
    unsigned long long s;
    explicit Rng(unsigned long long seed) : s(seed * 0x9E3779B97F4A7C15ull + 1) {}
    double uniform(double lo, double hi)
    {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return lo + (hi - lo) * (double)(s >> 11) * (1.0 / 9007199254740992.0);
    }
};

sc_params base(int ndim, int nx, int ny, int physics)
{
    sc_params p;
    std::memset(&p, 0, sizeof(p));
    p.ndim = ndim;
    p.nx[0] = nx;
    p.nx[1] = ndim == 2 ? ny : 1;
    p.mbc = 3;
    p.physics = physics;
    switch (physics) {
    case SC_PHYS_ADVECTION: p.meqn = p.mwaves = 1; break;
    case SC_PHYS_ACOUSTICS: p.meqn = p.mwaves = 2; break;
    case SC_PHYS_EULER:
        p.meqn = ndim == 1 ? 3 : 4;
        p.mwaves = 3;
        break;
    default:
        p.meqn = ndim == 1 ? 2 : 3;
        p.mwaves = ndim == 1 ? 2 : 3;
    }
    p.lim_type = 2;
    p.char_decomp = 2;
    p.time_integrator = 3;
    for (int k = 0; k < 4; k++) p.mthbc[k] = 1;
    p.xlower[0] = p.xlower[1] = 0.0;
    p.dx[0] = 1.0 / nx;
    p.dx[1] = ndim == 2 ? 1.0 / ny : 1.0;
    p.efix = -1;
    p.gamma = 1.4;
    p.grav = 1.0;
    p.u = 1.0;
    p.rho = 1.0;
    p.bulk = 4.0;
    p.precision = SC_FP64;
    for (int k = 0; k < 4; k++) p.mthlim[k] = 1 + k % 4;
    return p;
}

size_t cells(const sc_params &p)
{
    size_t n = (size_t)(p.nx[0] + 2 * p.mbc);
    if (p.ndim == 2) n *= (size_t)(p.nx[1] + 2 * p.mbc);
    return n;
}

// Physical cell states: a few superposed sine modes per field (smooth
// WENO regime), ranges shrunk by amp about their midpoints, with an optional
// constant block in 2D so that constant tiles occur.
std::vector<double> state(const sc_params &p, unsigned long long seed, bool flat_block = false, double amp = 0.5)
{
    Rng r(seed);
    const size_t n = cells(p);
    const int LX = p.nx[0] + 2 * p.mbc, LY = p.ndim == 2 ? p.nx[1] + 2 * p.mbc : 1;
    // field k: values in [0, 1]
    auto field = [&]() {
        std::vector<double> f(n, 0.0);
        for (int mode = 0; mode < 3; mode++) {
            const double kx = r.uniform(1, 4), ky = r.uniform(1, 4), ph = r.uniform(0, 6.283185307179586);
            for (int j = 0; j < LY; j++)
                for (int i = 0; i < LX; i++) {
                    const double x = LX > 1 ? (double)i / (LX - 1) : 0.0, y = LY > 1 ? (double)j / (LY - 1) : 0.0;
                    f[(size_t)i + (size_t)LX * j] += std::sin(6.283185307179586 * (kx * x + (LY > 1 ? ky * y : 0.0)) + ph);
                }
        }
        double lo = 1e300, hi = -1e300;
        for (double v : f) {
            lo = std::fmin(lo, v);
            hi = std::fmax(hi, v);
        }
        for (double &v : f) v = (v - lo) / std::fmax(hi - lo, 1e-300);
        return f;
    };
    auto range = [&](const std::vector<double> &f, size_t k, double lo, double hi) {
        const double mid = 0.5 * (lo + hi);
        lo = mid - amp * (mid - lo);
        hi = mid + amp * (hi - mid);
        const bool flat = flat_block && p.ndim == 2 && (int)(k % LX) < LX / 2 && (int)(k / LX) < LY / 2;
        return flat ? mid : lo + (hi - lo) * f[k];
    };
    std::vector<double> q(n * p.meqn);
    std::vector<std::vector<double>> F;
    for (int m = 0; m < 4; m++) F.push_back(field());
    for (size_t k = 0; k < n; k++) {
        double v[4];
        switch (p.physics) {
        case SC_PHYS_ADVECTION: v[0] = range(F[0], k, -1.0, 2.0); break;
        case SC_PHYS_ACOUSTICS:
            v[0] = range(F[0], k, -1.0, 1.0);
            v[1] = range(F[1], k, -1.0, 1.0);
            break;
        case SC_PHYS_EULER: {
            const double rho = range(F[0], k, 0.3, 2.0), u = range(F[1], k, -1.5, 1.5), pr = range(F[2], k, 0.3, 2.0);
            if (p.meqn == 3) {
                v[0] = rho;
                v[1] = rho * u;
                v[2] = pr / (p.gamma - 1.0) + 0.5 * rho * u * u;
            } else {
                const double w = range(F[3], k, -1.5, 1.5);
                v[0] = rho;
                v[1] = rho * u;
                v[2] = rho * w;
                v[3] = pr / (p.gamma - 1.0) + 0.5 * rho * (u * u + w * w);
            }
            break;
        }
        default: {
            const double h = range(F[0], k, 0.3, 2.0);
            v[0] = h;
            v[1] = h * range(F[1], k, -1.5, 1.5);
            if (p.meqn == 3) v[2] = h * range(F[2], k, -1.5, 1.5);
        }
        }
        for (int m = 0; m < p.meqn; m++) q[k + n * m] = v[m];
    }
    return q;
}

// max |a - b| over the interior / max |b| over the interior (infinite when
// either holds a non-finite value)
double rel_linf(const sc_params &p, const std::vector<double> &a, const std::vector<double> &b)
{
    const int mbc = p.mbc, LX = p.nx[0] + 2 * mbc, ny = p.ndim == 2 ? p.nx[1] : 1;
    const size_t n = cells(p);
    double d = 0.0, m = 0.0;
    for (int c = 0; c < p.meqn; c++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < p.nx[0]; i++) {
                const size_t k = (size_t)(i + mbc) + (p.ndim == 2 ? (size_t)LX * (size_t)(j + mbc) : 0) + n * c;
                if (!std::isfinite(a[k]) || !std::isfinite(b[k])) return INFINITY;
                d = std::fmax(d, std::fabs(a[k] - b[k]));
                m = std::fmax(m, std::fabs(b[k]));
            }
    return d / (m > 0 ? m : 1.0);
}

double rel_l1(const sc_params &p, const std::vector<double> &a, const std::vector<double> &b)
{
    const int mbc = p.mbc, LX = p.nx[0] + 2 * mbc, ny = p.ndim == 2 ? p.nx[1] : 1;
    const size_t n = cells(p);
    double d = 0.0, m = 0.0;
    for (int c = 0; c < p.meqn; c++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < p.nx[0]; i++) {
                const size_t k = (size_t)(i + mbc) + (p.ndim == 2 ? (size_t)LX * (size_t)(j + mbc) : 0) + n * c;
                if (!std::isfinite(a[k]) || !std::isfinite(b[k])) return INFINITY;
                d += std::fabs(a[k] - b[k]);
                m += std::fabs(b[k]);
            }
    return d / (m > 0 ? m : 1.0);
}

int g_fail = 0;

bool check(bool ok, const char *what, const std::string &name, double v, double tol)
{
    if (!ok) {
        std::printf("  %s: %s %.3e > %.1e\n", name.c_str(), what, v, tol);
        g_fail++;
    }
    return ok;
}

bool rc_ok(int rc, const char *call)
{
    if (rc != SC_OK) {
        std::printf("  %s failed (%d): %s\n", call, rc, sc_last_error());
        g_fail++;
        return false;
    }
    return true;
}

// One time step through sc_ts against orc_ts.
bool one_step(const std::string &name, sc_params p, unsigned long long seed, int tile = -1, bool flat = false,
              const std::vector<double> *aux = nullptr)
{
    std::vector<double> q = state(p, seed, flat);
    const double dt = 0.2 * (p.ndim == 2 ? std::fmin(p.dx[0], p.dx[1]) : p.dx[0]) / 3.0;
    sc_solver *s = nullptr;
    if (!rc_ok(sc_create(&p, &s), "sc_create")) return false;
    bool ok = true;
    if (tile >= 0) ok = rc_ok(sc_set_tile(s, tile), "sc_set_tile");
    if (ok && aux) ok = rc_ok(sc_set_aux(s, aux->data()), "sc_set_aux");
    ok = ok && rc_ok(sc_set_q(s, q.data()), "sc_set_q") && rc_ok(sc_save_qold(s), "sc_save_qold");
    double cfl = 0.0;
    ok = ok && rc_ok(sc_ts(s, dt, 0.0, &cfl), "sc_ts");
    std::vector<double> qd(q.size());
    ok = ok && rc_ok(sc_get_q(s, qd.data()), "sc_get_q");
    sc_destroy(s);
    if (!ok) return false;
    std::vector<double> qo = q, qold = q, r1(q.size()), r2(q.size());
    long long nev = 0;
    orc_set_aux(aux ? aux->data() : nullptr);
    const double cflo = orc_ts(&p, qo.data(), qold.data(), r1.data(), r2.data(), dt, 0.0, &nev);
    orc_set_aux(nullptr);
    if (p.precision == SC_FP32) {
        const double e = rel_l1(p, qd, qo);
        return check(e <= 1e-5, "relative L1 vs oracle", name, e, 1e-5) &&
               check(std::fabs(cfl - cflo) <= 1e-5 * cflo, "cfl", name, std::fabs(cfl - cflo), 1e-5 * cflo);
    }
    const double e = rel_linf(p, qd, qo);
    if (std::getenv("EMU_VERBOSE")) std::printf("  cfl %.17g oracle %.17g, q error %.3e\n", cfl, cflo, e);
    return check(e <= 1e-12, "relative Linf vs oracle", name, e, 1e-12) &&
           check(std::fabs(cfl - cflo) <= 1e-12 * cflo, "cfl", name, std::fabs(cfl - cflo), 1e-12 * cflo);
}

// The whole time loop through sc_advance against orc_sharpclaw.
bool advance(const std::string &name, sc_params p, unsigned long long seed, int steps, double tol)
{
    std::vector<double> q = state(p, seed);
    sc_run_ctl c = {1e99, 1.0, 0.9, steps, 1};
    const double dt0 = 0.1 * p.dx[0];
    sc_solver *s = nullptr;
    if (!rc_ok(sc_create(&p, &s), "sc_create")) return false;
    double dt = dt0;
    sc_run_stats st;
    bool ok = rc_ok(sc_set_q(s, q.data()), "sc_set_q") && rc_ok(sc_advance(s, &c, 0.0, 1e9, &dt, &st), "sc_advance");
    std::vector<double> qd(q.size());
    ok = ok && rc_ok(sc_get_q(s, qd.data()), "sc_get_q");
    sc_destroy(s);
    if (!ok) return false;
    std::vector<double> qo = q;
    double dto = dt0;
    sc_run_stats so;
    orc_sharpclaw(&p, &c, qo.data(), 0.0, 1e9, &dto, &so);
    const double e = p.precision == SC_FP32 ? rel_l1(p, qd, qo) : rel_linf(p, qd, qo);
    return check(e <= tol, "relative error vs oracle", name, e, tol) &&
           check(st.steps_taken == so.steps_taken, "steps taken differ", name, st.steps_taken, so.steps_taken);
}

bool noalloc(const std::string &name, sc_params p, unsigned long long seed)
{
    std::vector<double> q = state(p, seed);
    const double dt = 0.05 * p.dx[0];
    // the solver's step
    sc_solver *s = nullptr;
    if (!rc_ok(sc_create(&p, &s), "sc_create")) return false;
    std::vector<double> q1 = q, dq1(q.size());
    double cfl1 = 0.0;
    bool ok = rc_ok(sc_set_q(s, q1.data()), "sc_set_q") && rc_ok(sc_step(s, dt, 0.0, dq1.data(), &cfl1), "sc_step");
    sc_destroy(s);
    if (!ok) return false;
    // the allocation-free entry point on "device" arrays (host memory here)
    std::vector<double> q2 = q, dq2(q.size(), 1.0);
    double cfl2 = -1.0;
    if (!rc_ok(sc_step_noalloc(&p, q2.data(), dq2.data(), dt, 0.0, &cfl2, nullptr), "sc_step_noalloc")) return false;
    double d = 0.0;
    for (size_t k = 0; k < q.size(); k++) d = std::fmax(d, std::fabs(dq1[k] - dq2[k]));
    return check(d == 0.0, "dq differs from sc_step", name, d, 0.0) &&
           check(cfl1 == cfl2, "cfl differs from sc_step", name, std::fabs(cfl1 - cfl2), 0.0);
}

bool batch(const std::string &name, sc_params p, const std::vector<int> &mx, double tend, double tol)
{
    const int ns = (int)mx.size();
    std::vector<double> xl(ns, 0.0), xu(ns, 1.0);
    sc_batch *b = nullptr;
    if (!rc_ok(sc_batch_create(&p, ns, mx.data(), xl.data(), xu.data(), &b), "sc_batch_create")) return false;
    std::vector<std::vector<double>> q0(ns);
    bool ok = true;
    for (int i = 0; i < ns && ok; i++) {
        sc_params pi = p;
        pi.nx[0] = mx[i];
        pi.dx[0] = 1.0 / mx[i];
        q0[i] = state(pi, 100 + i);
        ok = rc_ok(sc_batch_set_q(b, i, q0[i].data()), "sc_batch_set_q");
    }
    sc_run_ctl c = {1e99, 1.0, 0.9, 100000, 1};
    std::vector<sc_batch_result> res(ns);
    double ms = 0.0;
    ok = ok && rc_ok(sc_batch_run(b, &c, 0.0, tend, 1e-4, res.data(), &ms), "sc_batch_run");
    for (int i = 0; i < ns && ok; i++) {
        sc_params pi = p;
        pi.nx[0] = mx[i];
        pi.dx[0] = 1.0 / mx[i];
        std::vector<double> qd(q0[i].size()), qo = q0[i];
        ok = rc_ok(sc_batch_get_q(b, i, qd.data()), "sc_batch_get_q");
        double dt = 1e-4;
        sc_run_stats so;
        orc_sharpclaw(&pi, &c, qo.data(), 0.0, tend, &dt, &so);
        const double e = p.precision == SC_FP32 ? rel_l1(pi, qd, qo) : rel_linf(pi, qd, qo);
        ok = ok && check(e <= tol, "relative error vs oracle", name + " solve " + std::to_string(i), e, tol);
        ok = ok && check(res[i].steps_taken == so.steps_taken, "steps taken differ", name, res[i].steps_taken,
                         so.steps_taken);
    }
    sc_batch_destroy(b);
    return ok;
}

bool serve(const std::string &name)
{
    sc_server *sv = nullptr;
    if (!rc_ok(sc_server_create(0, (size_t)256 << 20, 8, &sv), "sc_server_create")) return false;
    std::vector<sc_params> ps;
    ps.push_back(base(2, 24, 20, SC_PHYS_SHALLOW_WATER));
    ps.push_back(base(2, 40, 33, SC_PHYS_SHALLOW_WATER));
    ps.push_back(base(1, 300, 1, SC_PHYS_EULER));
    ps.push_back(base(1, 90, 1, SC_PHYS_EULER));
    ps.push_back(base(1, 260, 1, SC_PHYS_EULER));
    ps.push_back(base(1, 120, 1, SC_PHYS_SHALLOW_WATER));
    ps[1].precision = SC_FP32;
    ps[3].precision = SC_FP32;
    ps[4].char_decomp = 1;  // weno5_wave
    ps[5].lim_type = 1;     // tvd2_char
    ps[5].char_decomp = 2;
    sc_run_ctl c = {1e99, 1.0, 0.9, 100000, 1};
    const double tend = 0.02;
    std::vector<long long> ids(ps.size());
    std::vector<std::vector<double>> q0(ps.size());
    bool ok = true;
    for (size_t i = 0; i < ps.size() && ok; i++) {
        q0[i] = state(ps[i], 300 + i);
        ok = rc_ok(sc_server_submit(sv, &ps[i], &c, 0.0, tend, 1e-4, q0[i].data(), &ids[i]), "sc_server_submit");
    }
    ok = ok && rc_ok(sc_server_drain(sv, 600000.0), "sc_server_drain");
    for (size_t i = 0; i < ps.size() && ok; i++) {
        std::vector<double> qd(q0[i].size()), qo = q0[i];
        sc_request_stats st;
        ok = rc_ok(sc_server_result(sv, ids[i], qd.data(), &st), "sc_server_result");
        double dt = 1e-4;
        sc_run_stats so;
        orc_sharpclaw(&ps[i], &c, qo.data(), 0.0, tend, &dt, &so);
        const bool f32 = ps[i].precision == SC_FP32;
        const double e = f32 ? rel_l1(ps[i], qd, qo) : rel_linf(ps[i], qd, qo);
        const double tol = f32 ? 1e-5 : 1e-10;
        ok = ok && check(e <= tol, "relative error vs oracle", name + " request " + std::to_string(i), e, tol);
    }
    sc_server_destroy(sv);
    return ok;
}

// The same batch with the largest solves split over several workgroups
// (the tuner's choice) and with every solve on one: bitwise equal.
bool batch_split(const std::string &name, sc_params p, const std::vector<int> &mx, double tend)
{
    const int ns = (int)mx.size();
    std::vector<double> xl(ns, 0.0), xu(ns, 1.0);
    std::vector<std::vector<double>> out[2];
    std::vector<int> parts(ns);
    int split = 0;
    for (int run = 0; run < 2; run++) {
        if (run == 0) setenv("SC_BATCH_PARTS", "1", 1);
        sc_batch *b = nullptr;
        if (!rc_ok(sc_batch_create(&p, ns, mx.data(), xl.data(), xu.data(), &b), "sc_batch_create")) return false;
        bool ok = true;
        for (int i = 0; i < ns && ok; i++) {
            sc_params pi = p;
            pi.nx[0] = mx[i];
            pi.dx[0] = 1.0 / mx[i];
            std::vector<double> q = state(pi, 500 + i);
            ok = rc_ok(sc_batch_set_q(b, i, q.data()), "sc_batch_set_q");
            out[run].push_back(std::vector<double>(q.size()));
        }
        sc_run_ctl c = {1e99, 1.0, 0.9, 100000, 1};
        std::vector<sc_batch_result> res(ns);
        double ms = 0.0;
        ok = ok && rc_ok(sc_batch_run(b, &c, 0.0, tend, 1e-4, res.data(), &ms), "sc_batch_run");
        ok = ok && rc_ok(sc_batch_parts(b, parts.data()), "sc_batch_parts");
        for (int i = 0; i < ns && ok; i++) {
            ok = rc_ok(sc_batch_get_q(b, i, out[run][i].data()), "sc_batch_get_q");
            if (run == 1 && parts[i] > 1) split++;
            if (run == 0 && parts[i] != 1) ok = check(false, "split although SC_BATCH_PARTS=1", name, parts[i], 1);
        }
        sc_batch_destroy(b);
        unsetenv("SC_BATCH_PARTS");
        if (!ok) return false;
    }
    bool ok = check(split > 0, "no solve was split", name, 0, 1);
    for (int i = 0; i < ns; i++)
        ok = ok && check(std::memcmp(out[0][i].data(), out[1][i].data(), out[0][i].size() * sizeof(double)) == 0,
                         "split and unsplit results differ", name + " solve " + std::to_string(i), 1, 0);
    return ok;
}

// sc_advance: the device-resident loop (1D: the grid split over its
// chunks; 2D: stage launches driven by the control block, several rounds)
// against the host-driven loop (SC_HOST_LOOP=1), bitwise, and against the
// oracle.
bool loop1d(const std::string &name, sc_params p, unsigned long long seed, double tend, double tol)
{
    std::vector<double> q = state(p, seed);
    sc_run_ctl c = {1e99, 1.0, 0.9, 100000, 1};
    std::vector<double> out[2];
    sc_run_stats st[2];
    for (int run = 0; run < 2; run++) {
        if (run == 0) setenv("SC_HOST_LOOP", "1", 1);
        sc_solver *s = nullptr;
        bool ok = rc_ok(sc_create(&p, &s), "sc_create");
        unsetenv("SC_HOST_LOOP");
        if (!ok) return false;
        double dt = 0.1 * p.dx[0];
        out[run].resize(q.size());
        ok = rc_ok(sc_set_q(s, q.data()), "sc_set_q") && rc_ok(sc_advance(s, &c, 0.0, tend, &dt, &st[run]), "sc_advance") &&
             rc_ok(sc_get_q(s, out[run].data()), "sc_get_q");
        sc_destroy(s);
        if (!ok) return false;
    }
    std::vector<double> qo = q;
    double dto = 0.1 * p.dx[0];
    sc_run_stats so;
    orc_sharpclaw(&p, &c, qo.data(), 0.0, tend, &dto, &so);
    const double e = p.precision == SC_FP32 ? rel_l1(p, out[1], qo) : rel_linf(p, out[1], qo);
    return check(std::memcmp(out[0].data(), out[1].data(), q.size() * sizeof(double)) == 0,
                 "device and host loops differ", name, 1, 0) &&
           check(st[0].steps_taken == st[1].steps_taken && st[0].dt_last == st[1].dt_last &&
                     st[0].cfl_maxused == st[1].cfl_maxused && st[0].stage_evals == st[1].stage_evals,
                 "loop statistics differ", name, st[1].steps_taken, st[0].steps_taken) &&
           check(e <= tol, "relative error vs oracle", name, e, tol) &&
           check(st[1].steps_taken == so.steps_taken, "steps taken differ", name, st[1].steps_taken, so.steps_taken);
}

// Steps of a 2D grid with a constant block (the deep-constant tiles of a
// step's second and third stages, which skip staging): sc_advance with the
// adaptive tile order (device loop, deep-constant tiles) against the band
// order (SC_TILE_ORDER=0: host loop, no tile flags), bitwise, and against
// the oracle.
bool deep2d(const std::string &name, sc_params p, unsigned long long seed, int steps, double tol)
{
    std::vector<double> q = state(p, seed, true);
    sc_run_ctl c = {1e99, 1.0, 0.9, steps, 1};
    std::vector<double> out[2];
    sc_run_stats st[2];
    for (int run = 0; run < 2; run++) {
        if (run == 1) setenv("SC_TILE_ORDER", "0", 1);
        sc_solver *s = nullptr;
        bool ok = rc_ok(sc_create(&p, &s), "sc_create");
        unsetenv("SC_TILE_ORDER");
        if (!ok) return false;
        double dt = 0.1 * p.dx[0];
        out[run].resize(q.size());
        ok = rc_ok(sc_set_q(s, q.data()), "sc_set_q") && rc_ok(sc_advance(s, &c, 0.0, 1e9, &dt, &st[run]), "sc_advance") &&
             rc_ok(sc_get_q(s, out[run].data()), "sc_get_q");
        sc_destroy(s);
        if (!ok) return false;
    }
    std::vector<double> qo = q;
    double dto = 0.1 * p.dx[0];
    sc_run_stats so;
    orc_sharpclaw(&p, &c, qo.data(), 0.0, 1e9, &dto, &so);
    const double e = rel_linf(p, out[0], qo);
    return check(std::memcmp(out[0].data(), out[1].data(), q.size() * sizeof(double)) == 0,
                 "adaptive order (deep-constant tiles) and band order differ", name, 1, 0) &&
           check(st[0].steps_taken == st[1].steps_taken && st[0].dt_last == st[1].dt_last,
                 "loop statistics differ", name, st[1].steps_taken, st[0].steps_taken) &&
           check(e <= tol, "relative error vs oracle", name, e, tol) &&
           check(st[0].steps_taken == so.steps_taken, "steps taken differ", name, st[0].steps_taken, so.steps_taken);
}

struct Scenario {
    std::string name;
    std::function<bool()> run;
};

std::vector<Scenario> scenarios()
{
    std::vector<Scenario> v;
    // 2D fused stage: every tile shape on a grid with partial tiles, both tile orders
    for (int order = 0; order < 2; order++)
        for (int tile = 0; tile < 4; tile++) {
            std::string n = "stage2d euler char tile " + std::to_string(tile) + " order " + std::to_string(order);
            v.push_back({n, [=]() {
                             setenv("SC_TILE_ORDER", order ? "1" : "0", 1);
                             sc_params p = base(2, 37, 29, SC_PHYS_EULER);
                             p.mthbc[0] = p.mthbc[1] = 3;
                             bool ok = one_step(n, p, 11 + tile, tile, true);
                             unsetenv("SC_TILE_ORDER");
                             return ok;
                         }});
        }
    v.push_back({"stage2d shallow water weno5 periodic", []() {
                     sc_params p = base(2, 33, 40, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 0;
                     for (int k = 0; k < 4; k++) p.mthbc[k] = 2;
                     p.time_integrator = 4;
                     return one_step("stage2d shallow water weno5 periodic", p, 21, 1);
                 }});
    v.push_back({"stage2d euler fp32", []() {
                     sc_params p = base(2, 30, 35, SC_PHYS_EULER);
                     p.precision = SC_FP32;
                     return one_step("stage2d euler fp32", p, 23, 3, true);
                 }});
    v.push_back({"stage2d swirl advection", []() {
                     sc_params p = base(2, 24, 24, SC_PHYS_ADVECTION);
                     p.maux = 2;
                     p.char_decomp = 0;
                     std::vector<double> aux(cells(p) * 2);
                     orc_swirl_setaux(&p, aux.data());
                     return one_step("stage2d swirl advection", p, 25, 0, false, &aux);
                 }});
    // capacity on 2D grids (aux(:,:,mcapa)), partial edge tiles, every
    // boundary option on some side; full path on a state with a flat block.
    // The ghost cells of side k (and its corners) get the smallest kappa, so
    // the CFL comes from that side's ghost row or column.
    const int capbc[4][4] = {{1, 3, 2, 2}, {2, 2, 3, 1}, {3, 3, 3, 3}, {3, 1, 1, 3}};
    for (int k = 0; k < 4; k++) {
        std::string n = "stage2d capacity ghost side " + std::to_string(k);
        v.push_back({n, [=]() {
                         sc_params p = base(2, 37 + k, 29 + 5 * k, k == 1 ? SC_PHYS_SHALLOW_WATER : SC_PHYS_EULER);
                         for (int b = 0; b < 4; b++) p.mthbc[b] = capbc[k][b];
                         p.maux = 2;
                         p.mcapa = 2;
                         p.char_decomp = k == 2 ? 0 : 2;
                         Rng r(900 + k);
                         const int LX = p.nx[0] + 6, LY = p.nx[1] + 6;
                         std::vector<double> aux(cells(p) * 2);
                         for (auto &x : aux) x = r.uniform(1.0, 2.0);
                         for (int j = 0; j < LY; j++)
                             for (int i = 0; i < LX; i++) {
                                 const int gi = i - 2, gj = j - 2;  // cell index 1..n inside
                                 const bool side = k == 0 ? gi == 0 : k == 1 ? gi == p.nx[0] + 1
                                                 : k == 2 ? gj == 0 : gj == p.nx[1] + 1;
                                 if (side) aux[(size_t)LX * LY + i + (size_t)LX * j] = r.uniform(0.2, 0.3);
                             }
                         return one_step(n, p, 27 + k, -1, k == 0, &aux);
                     }});
    }
    // 2D slice sweeps (reconstructions that read neighbouring interfaces)
    const int sweeps[3][2] = {{2, 1}, {2, 3}, {1, 1}};
    for (auto &sw : sweeps) {
        std::string n = "sweep2d euler lim " + std::to_string(sw[0]) + " char " + std::to_string(sw[1]);
        v.push_back({n, [=]() {
                         sc_params p = base(2, 26, 21, SC_PHYS_EULER);
                         p.lim_type = sw[0];
                         p.char_decomp = sw[1];
                         return one_step(n, p, 31 + sw[1]);
                     }});
    }
    // 1D fused stage: several chunks per grid, every physics
    const int phys[4] = {SC_PHYS_EULER, SC_PHYS_SHALLOW_WATER, SC_PHYS_ACOUSTICS, SC_PHYS_ADVECTION};
    for (int ph : phys) {
        std::string n = "stage1d physics " + std::to_string(ph);
        v.push_back({n, [=]() {
                         sc_params p = base(1, 700, 1, ph);
                         p.char_decomp = ph == SC_PHYS_EULER ? 2 : 0;
                         p.time_integrator = 4;
                         if (ph != SC_PHYS_ADVECTION) p.mthbc[0] = 3;
                         return one_step(n, p, 41 + ph);
                     }});
    }
    v.push_back({"stage1d shallow water tfluct", []() {
                     sc_params p = base(1, 300, 1, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 0;
                     p.tfluct_solver = 1;
                     return one_step("stage1d shallow water tfluct", p, 51);
                 }});
    const int recon1d[4][2] = {{2, 1}, {2, 3}, {1, 1}, {1, 2}};
    for (auto &rc : recon1d) {
        std::string n = "stage1d euler lim " + std::to_string(rc[0]) + " char " + std::to_string(rc[1]);
        v.push_back({n, [=]() {
                         sc_params p = base(1, 400, 1, SC_PHYS_EULER);
                         p.lim_type = rc[0];
                         p.char_decomp = rc[1];
                         return one_step(n, p, 61 + rc[1]);
                     }});
    }
    v.push_back({"stage1d euler fp32", []() {
                     sc_params p = base(1, 500, 1, SC_PHYS_EULER);
                     p.precision = SC_FP32;
                     return one_step("stage1d euler fp32", p, 71);
                 }});
    // time loops
    v.push_back({"advance2d euler ssp104", []() {
                     sc_params p = base(2, 40, 34, SC_PHYS_EULER);
                     p.time_integrator = 4;
                     return advance("advance2d euler ssp104", p, 81, 4, 1e-11);
                 }});
    v.push_back({"advance1d euler ssp33", []() {
                     sc_params p = base(1, 600, 1, SC_PHYS_EULER);
                     return advance("advance1d euler ssp33", p, 83, 6, 1e-11);
                 }});
    v.push_back({"loop2d euler ssp33", []() {
                     sc_params p = base(2, 48, 40, SC_PHYS_EULER);
                     return loop1d("loop2d euler ssp33", p, 85, 0.15, 1e-11);
                 }});
    v.push_back({"loop2d shallow water fp32 ssp22", []() {
                     sc_params p = base(2, 36, 52, SC_PHYS_SHALLOW_WATER);
                     p.precision = SC_FP32;
                     p.time_integrator = 1;
                     return loop1d("loop2d shallow water fp32 ssp22", p, 87, 0.05, 1e-5);
                 }});
    v.push_back({"deep2d euler ssp33", []() {
                     return deep2d("deep2d euler ssp33", base(2, 128, 120, SC_PHYS_EULER), 89, 4, 1e-11);
                 }});
    v.push_back({"deep2d shallow water ssp104", []() {
                     sc_params p = base(2, 112, 128, SC_PHYS_SHALLOW_WATER);
                     p.time_integrator = 4;
                     return deep2d("deep2d shallow water ssp104", p, 91, 2, 1e-11);
                 }});
    v.push_back({"noalloc euler2d", []() { return noalloc("noalloc euler2d", base(2, 45, 38, SC_PHYS_EULER), 91); }});
    v.push_back({"noalloc sw1d", []() {
                     sc_params p = base(1, 900, 1, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 0;
                     return noalloc("noalloc sw1d", p, 93);
                 }});
    v.push_back({"batch euler fp64", []() {
                     return batch("batch euler fp64", base(1, 64, 1, SC_PHYS_EULER), {64, 100, 257, 513, 700, 90, 33},
                                  0.01, 1e-10);
                 }});
    v.push_back({"batch shallow water fp32", []() {
                     sc_params p = base(1, 64, 1, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 0;
                     p.precision = SC_FP32;
                     return batch("batch shallow water fp32", p, {48, 130, 600, 77}, 0.01, 1e-5);
                 }});
    v.push_back({"batch split euler fp64", []() {
                     return batch_split("batch split euler fp64", base(1, 64, 1, SC_PHYS_EULER), {1500, 64, 90, 33},
                                        0.01);
                 }});
    v.push_back({"batch split shallow water fp32 ssp104", []() {
                     sc_params p = base(1, 64, 1, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 0;
                     p.precision = SC_FP32;
                     p.time_integrator = 4;
                     return batch_split("batch split shallow water fp32 ssp104", p, {2000, 700, 50}, 0.005);
                 }});
    // the reconstructions with neighbour data in batches, split solves and
    // the 1D device loop (flux1.f90:65-117)
    v.push_back({"batch euler weno5_wave fp64", []() {
                     sc_params p = base(1, 64, 1, SC_PHYS_EULER);
                     p.char_decomp = 1;
                     return batch("batch euler weno5_wave fp64", p, {64, 300, 257, 33}, 0.01, 1e-10);
                 }});
    v.push_back({"batch shallow water weno5_trans fp32", []() {
                     sc_params p = base(1, 64, 1, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 3;
                     p.precision = SC_FP32;
                     return batch("batch shallow water weno5_trans fp32", p, {48, 400, 90}, 0.01, 1e-5);
                 }});
    v.push_back({"batch acoustics tvd2_wave ssp22", []() {
                     sc_params p = base(1, 64, 1, SC_PHYS_ACOUSTICS);
                     p.lim_type = 1;
                     p.char_decomp = 1;
                     p.time_integrator = 1;
                     p.mthbc[0] = 3;
                     return batch("batch acoustics tvd2_wave ssp22", p, {200, 64, 513}, 0.02, 1e-10);
                 }});
    v.push_back({"batch split euler tvd2_char", []() {
                     sc_params p = base(1, 64, 1, SC_PHYS_EULER);
                     p.lim_type = 1;
                     p.char_decomp = 2;
                     return batch_split("batch split euler tvd2_char", p, {1400, 64, 90}, 0.01);
                 }});
    v.push_back({"loop1d euler weno5_trans", []() {
                     sc_params p = base(1, 900, 1, SC_PHYS_EULER);
                     p.char_decomp = 3;
                     p.mthbc[1] = 3;
                     return loop1d("loop1d euler weno5_trans", p, 607, 0.005, 1e-10);
                 }});
    v.push_back({"loop1d euler ssp104", []() {
                     sc_params p = base(1, 1600, 1, SC_PHYS_EULER);
                     p.time_integrator = 4;
                     p.mthbc[0] = 3;
                     return loop1d("loop1d euler ssp104", p, 601, 0.004, 1e-10);
                 }});
    v.push_back({"loop1d shallow water fp32", []() {
                     sc_params p = base(1, 700, 1, SC_PHYS_SHALLOW_WATER);
                     p.char_decomp = 0;
                     p.precision = SC_FP32;
                     return loop1d("loop1d shallow water fp32", p, 603, 0.01, 1e-5);
                 }});
    // parts whose widened ranges reach a boundary; parts owning more
    // than one chunk per stage (MAX_PARTS reached); periodic boundaries
    // (a barrier after every stage)
    v.push_back({"loop1d euler ssp104 narrow parts", []() {
                     sc_params p = base(1, 200, 1, SC_PHYS_EULER);
                     p.time_integrator = 4;
                     p.mthbc[1] = 3;
                     return loop1d("loop1d euler ssp104 narrow parts", p, 611, 0.01, 1e-10);
                 }});
    v.push_back({"loop1d euler ssp104 multi-chunk parts", []() {
                     sc_params p = base(1, 4000, 1, SC_PHYS_EULER);
                     p.time_integrator = 4;
                     return loop1d("loop1d euler ssp104 multi-chunk parts", p, 613, 0.0004, 1e-10);
                 }});
    v.push_back({"loop1d euler periodic", []() {
                     sc_params p = base(1, 600, 1, SC_PHYS_EULER);
                     p.mthbc[0] = p.mthbc[1] = 2;
                     return loop1d("loop1d euler periodic", p, 615, 0.005, 1e-10);
                 }});
    v.push_back({"serve mixed", []() { return serve("serve mixed"); }});
    return v;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *only = argc > 1 ? argv[1] : nullptr;
    int ran = 0;
    for (auto &s : scenarios()) {
        if (only && s.name.find(only) == std::string::npos) continue;
        const int before = g_fail;
        const bool ok = s.run() && g_fail == before;
        std::printf("%s %s\n", ok ? "PASS" : "FAIL", s.name.c_str());
        std::fflush(stdout);
        if (!ok && g_fail == before) g_fail++;
        ran++;
    }
    std::printf("%d scenarios, %d failures\n", ran, g_fail);
    return g_fail == 0 && ran > 0 ? 0 : 1;
}
