// TEST INFRASTRUCTURE ONLY -- the CPU oracle for the FDTD hot path.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
// only to check the engine's results against, or as the timed CPU baseline.  The product
// (rfnetwork_amd) never calls it.
//
// It restates, as plain loops, the threaded C++ solver of rfnetwork:
//   rfnetwork/core/src/solver.cpp:239-349   solver_init_fields  (shape checks, PEC zeroing of the
//                                            last x plane of Cy, Cz and Dx -- mutates the caller arrays)
//   rfnetwork/core/src/solver.cpp:351-463   solver_init_monitors (CPU-layout strides and shifts)
//   rfnetwork/core/src/solver.cpp:465-532   solver_init_probes   (global Yee index -> slab index)
//   rfnetwork/core/src/solver.cpp:534-674   solver_run / solver_controller (x slabs, controller
//                                            thread, two condition-variable barriers per step)
//   rfnetwork/core/src/solver.cpp:676-1165  solver_thread (E loop, E probes, barrier, H loop,
//                                            H probes, barrier, monitors)
//   rfnetwork/core/src/solver.cpp:1168-1225 update_monitor / update_phasor_monitor
// The reference expresses the per-plane updates with Eigen row-major maps; Eigen's cwiseProduct,
// + and - are single IEEE operations per element, so the loops below evaluate the same expression
// trees in the same order.  Build with -ffp-contract=off so that no multiply-add is fused (the
// reference is built for baseline x86-64, which has no FMA).
//
// Layout (the reference's CPU packing, solver.py:1473-1523 with end=-1, hs=0):
//   E coefficients  Cx (Nx, Ny-1, Nz-1), Cy (Nx, Ny, Nz-1), Cz (Nx, Ny-1, Nz)
//   H coefficients  Dx (Nx, Ny, Nz),     Dy (Nx, Ny+1, Nz), Dz (Nx, Ny, Nz+1)
//   per-slab planes ex (Ny+1)(Nz+1), ey Ny(Nz+1), ez (Ny+1)Nz, hx NyNz, hy (Ny+1)Nz, hz Ny(Nz+1)
//   slab plane x holds Ex, Hy, Hz at global x_start+x and Ey, Ez, Hx at global x_start+x+1.
//
// Differences from the reference, all on inputs where the reference crashes or is undefined:
// errors are returned as codes with a message (the reference throws through the CPython API and
// aborts); n_threads == 0 and n_threads > Nx are rejected (division by zero / zero-width slabs).
#include <atomic>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

enum { EX = 0, EY = 1, EZ = 2, HX = 3, HY = 4, HZ = 5 };
constexpr int kMaxMonitors = 50;   // solver.h:15
constexpr int kMaxProbes = 5000;   // solver.h:16
constexpr int kMaxThreads = 20;    // solver.h:17

thread_local std::string g_err;

struct OracleError {
    std::string msg;
};

}  // namespace

extern "C" {

// Probe record: values is R[>= Nt], source samples in, sampled total field out.
typedef struct oracle_probe {
    int32_t field;
    int32_t idx[3];  // global Yee indices (solver.py:1557-1575)
    int32_t is_source;
    int32_t reserved;
    void* values;
} oracle_probe;

// Monitor record (solver.py:1597-1620).  n_frequencies == 0 selects a time-domain monitor whose
// values are R[Nt/n_step+1, N1, N2]; otherwise values are complex<R>[Nf, N1, N2] and dtft_phase is
// complex<R>[Nt/n_step+1, Nf].
typedef struct oracle_monitor {
    int32_t axis;
    int32_t position;
    int32_t field;
    int32_t n_step;
    int32_t n_frequencies;
    int32_t reserved;
    void* values;
    const void* dtft_phase;
} oracle_monitor;

const char* oracle_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

template <typename R>
struct Coeffs {
    // E: Cx, Cy, Cz; H: Dx, Dy, Dz.  Names follow the coefficient dictionary of solver.py:1473-1518.
    R *Ca_ex_y, *Ca_ex_z, *Cb_ex_y, *Cb_ex_z;
    R *Ca_ey_z, *Ca_ey_x, *Cb_ey_z, *Cb_ey_x;
    R *Ca_ez_x, *Ca_ez_y, *Cb_ez_x, *Cb_ez_y;
    R *Da_hx_y, *Da_hx_z, *Db_hx_y1, *Db_hx_y2, *Db_hx_z1, *Db_hx_z2;
    R *Da_hy_z, *Da_hy_x, *Db_hy_z1, *Db_hy_z2, *Db_hy_x1, *Db_hy_x2;
    R *Da_hz_x, *Da_hz_y, *Db_hz_x1, *Db_hz_x2, *Db_hz_y1, *Db_hz_y2;
};

struct ProbeSlot {
    int field, x, y, z;  // shifted (slab-grid) indices
    bool is_source;
};

struct MonitorSlot {
    int field, axis, position, n_step, n_phasors;
    int N1, N2, row_stride, col_stride, yz_offset;
    void* values;
    const void* phase;
};

// Halo planes exchanged between neighbouring slabs (solver.h:109-114).
template <typename R>
struct Halo {
    R* hy = nullptr;  // first Hy plane of the slab (read by the previous slab's E update)
    R* hz = nullptr;
    R* ey = nullptr;  // last Ey plane of the slab (read by the next slab's H update)
    R* ez = nullptr;
};

template <typename R>
class Oracle {
   public:
    int Nx, Ny, Nz, Nt;
    Coeffs<R> C{};
    std::vector<ProbeSlot> probes;
    std::vector<R*> probe_values;
    std::vector<MonitorSlot> monitors;

    // scratch pool carved from the caller's mem array (solver.cpp:194-227)
    R* pool_next = nullptr;
    uint64_t pool_left = 0;
    std::mutex mu;
    std::condition_variable cv_ctrl, cv_work;
    int init_count = 0, e_count = 0, h_count = 0;
    bool init_done = false, e_done = false, h_done = false;
    std::vector<Halo<R>> halo;
    R* const* fields_out = nullptr;  // optional 18 global Yee arrays written after the run
    std::vector<std::string> thread_errors;

    R* take(uint64_t n) {
        std::lock_guard<std::mutex> lk(mu);
        if (n > pool_left) throw OracleError{"Memory buffer has insufficent space."};
        R* p = pool_next;
        pool_next += n;
        pool_left -= n;
        return p;
    }

    void run(int n_threads, int update_interval);
    void controller(int n_threads, int update_interval);
    void worker(int x_start, int x_stop, int t);
};

// Progress bar of solver.cpp:50-76 (20 dots), printed by the controller thread.
void progress(int n, int Nt) {
    int pct = (n + 1) * 100 / Nt;
    int dots = pct / 5;
    std::string s = "\r|";
    s.append(dots, '.');
    s.append(20 - dots, ' ');
    s += "| " + std::to_string(pct) + "%  ";
    std::fputs(s.c_str(), stdout);
    std::fflush(stdout);
}

template <typename R>
void Oracle<R>::controller(int n_threads, int update_interval) {
    // solver.cpp:599-674: wait for every worker's allocation, then run 2 barriers per step.
    {
        std::unique_lock<std::mutex> lk(mu);
        cv_ctrl.wait(lk, [&] { return init_count == n_threads; });
        init_done = true;
        cv_work.notify_all();
    }
    if (update_interval) progress(0, Nt);
    for (int n = 0; n < Nt; ++n) {
        {
            std::unique_lock<std::mutex> lk(mu);
            h_count = 0;
            h_done = true;
            e_done = false;
            cv_work.notify_all();
            cv_ctrl.wait(lk, [&] { return e_count == n_threads; });
        }
        {
            std::unique_lock<std::mutex> lk(mu);
            e_count = 0;
            e_done = true;
            h_done = false;
            cv_work.notify_all();
            cv_ctrl.wait(lk, [&] { return h_count == n_threads; });
        }
        if (update_interval && (n % update_interval) == 0) progress(n, Nt);
    }
    {
        std::unique_lock<std::mutex> lk(mu);
        h_count = 0;
        h_done = true;
        e_done = false;
        cv_work.notify_all();
    }
    if (update_interval > 0) progress(Nt, Nt);
}

template <typename R>
void Oracle<R>::worker(int x_start, int x_stop, int t) {
    const int nx = x_stop - x_start;
    const int Nyp1 = Ny + 1, Nzp1 = Nz + 1, Nym1 = Ny - 1, Nzm1 = Nz - 1;
    const int ex_p = Nyp1 * Nzp1, ey_p = Ny * Nzp1, ez_p = Nyp1 * Nz;
    const int hx_p = Ny * Nz, hy_p = Nyp1 * Nz, hz_p = Ny * Nzp1;
    const int cx_p = Nym1 * Nzm1, cy_p = Ny * Nzm1, cz_p = Nym1 * Nz;
    const int dx_p = Ny * Nz, dy_p = Nyp1 * Nz, dz_p = Ny * Nzp1;

    R *ex_y, *ex_z, *ex, *ey_z, *ey_x, *ey, *ez_x, *ez_y, *ez;
    R *hx_y, *hx_z, *hx, *hy_z, *hy_x, *hy, *hz_x, *hz_y, *hz;
    std::vector<ProbeSlot*> e_probes, h_probes;
    std::vector<R*> e_vals, h_vals;
    std::vector<R*> p_tot, p_s1, p_s2;  // per probe in e_probes then h_probes order
    try {
        // solver.cpp:725-748: 18 slab arrays from the pool
        ex_y = take((uint64_t)nx * ex_p); ex_z = take((uint64_t)nx * ex_p); ex = take((uint64_t)nx * ex_p);
        ey_z = take((uint64_t)nx * ey_p); ey_x = take((uint64_t)nx * ey_p); ey = take((uint64_t)nx * ey_p);
        ez_x = take((uint64_t)nx * ez_p); ez_y = take((uint64_t)nx * ez_p); ez = take((uint64_t)nx * ez_p);
        hx_y = take((uint64_t)nx * hx_p); hx_z = take((uint64_t)nx * hx_p); hx = take((uint64_t)nx * hx_p);
        hy_z = take((uint64_t)nx * hy_p); hy_x = take((uint64_t)nx * hy_p); hy = take((uint64_t)nx * hy_p);
        hz_x = take((uint64_t)nx * hz_p); hz_y = take((uint64_t)nx * hz_p); hz = take((uint64_t)nx * hz_p);
    } catch (OracleError& e) {
        std::lock_guard<std::mutex> lk(mu);
        thread_errors.push_back(e.msg);
        ex_y = nullptr;
    }
    if (ex_y) {
        for (R* a : {ex_y, ex_z, ex}) std::memset(a, 0, sizeof(R) * (size_t)nx * ex_p);
        for (R* a : {ey_z, ey_x, ey}) std::memset(a, 0, sizeof(R) * (size_t)nx * ey_p);
        for (R* a : {ez_x, ez_y, ez}) std::memset(a, 0, sizeof(R) * (size_t)nx * ez_p);
        for (R* a : {hx_y, hx_z, hx}) std::memset(a, 0, sizeof(R) * (size_t)nx * hx_p);
        for (R* a : {hy_z, hy_x, hy}) std::memset(a, 0, sizeof(R) * (size_t)nx * hy_p);
        for (R* a : {hz_x, hz_y, hz}) std::memset(a, 0, sizeof(R) * (size_t)nx * hz_p);
        halo[t].hy = hy;
        halo[t].hz = hz;
        halo[t].ey = ey + (size_t)(nx - 1) * ey_p;
        halo[t].ez = ez + (size_t)(nx - 1) * ez_p;

        // solver.cpp:771-832: probes inside this slab, with their offsets in the slab arrays
        R* base[6] = {ex, ey, ez, hx, hy, hz};
        R* sp1[6] = {ex_y, ey_z, ez_x, hx_y, hy_z, hz_x};
        R* sp2[6] = {ex_z, ey_x, ez_y, hx_z, hy_x, hz_y};
        for (size_t i = 0; i < probes.size(); ++i) {
            ProbeSlot& p = probes[i];
            if (p.x < x_start || p.x >= x_stop) continue;
            if (p.field <= EZ) {
                e_probes.push_back(&p);
                e_vals.push_back(probe_values[i]);
            } else {
                h_probes.push_back(&p);
                h_vals.push_back(probe_values[i]);
            }
        }
        // the probe's component inside the slab arrays (total, first split, second split)
        auto offset_of = [&](const ProbeSlot& p, R** tot, R** s1, R** s2) {
            const int lx = p.x - x_start;
            size_t off;
            switch (p.field) {
                case EX: off = (size_t)lx * ex_p + (size_t)(p.y + 1) * Nzp1 + (p.z + 1); break;
                case EY: off = (size_t)lx * ey_p + (size_t)p.y * Nzp1 + (p.z + 1); break;
                case EZ: off = (size_t)lx * ez_p + (size_t)(p.y + 1) * Nz + p.z; break;
                case HX: off = (size_t)lx * hx_p + (size_t)p.y * Nz + p.z; break;
                case HY: off = (size_t)lx * hy_p + (size_t)(p.y + 1) * Nz + p.z; break;
                default: off = (size_t)lx * hz_p + (size_t)p.y * Nzp1 + (p.z + 1); break;
            }
            *tot = base[p.field] + off;
            *s1 = sp1[p.field] + off;
            *s2 = sp2[p.field] + off;
        };
        for (ProbeSlot* p : e_probes) {
            R *a, *b, *c;
            offset_of(*p, &a, &b, &c);
            p_tot.push_back(a); p_s1.push_back(b); p_s2.push_back(c);
        }
        for (ProbeSlot* p : h_probes) {
            R *a, *b, *c;
            offset_of(*p, &a, &b, &c);
            p_tot.push_back(a); p_s1.push_back(b); p_s2.push_back(c);
        }
    }

    bool ok;
    {
        std::unique_lock<std::mutex> lk(mu);
        ++init_count;
        cv_ctrl.notify_all();
        cv_work.wait(lk, [&] { return init_done; });
        ok = thread_errors.empty();
    }
    const size_t n_e = e_probes.size();

    for (int n = 0; n < Nt; ++n) {
        if (ok) {
            // ---- E half step, solver.cpp:851-960 ----
            for (int x = 0; x < nx; ++x) {
                R* exy = ex_y + (size_t)x * ex_p; R* exz = ex_z + (size_t)x * ex_p; R* ext = ex + (size_t)x * ex_p;
                R* eyz = ey_z + (size_t)x * ey_p; R* eyx = ey_x + (size_t)x * ey_p; R* eyt = ey + (size_t)x * ey_p;
                R* ezx = ez_x + (size_t)x * ez_p; R* ezy = ez_y + (size_t)x * ez_p; R* ezt = ez + (size_t)x * ez_p;
                const R* hxp = hx + (size_t)x * hx_p;
                const R* hyp = hy + (size_t)x * hy_p;
                const R* hzp = hz + (size_t)x * hz_p;
                const R* hy1 = (x < nx - 1) ? hy + (size_t)(x + 1) * hy_p : halo[t + 1].hy;
                const R* hz1 = (x < nx - 1) ? hz + (size_t)(x + 1) * hz_p : halo[t + 1].hz;
                const size_t xg = (size_t)(x_start + x);
                const R* ca_xy = C.Ca_ex_y + xg * cx_p; const R* ca_xz = C.Ca_ex_z + xg * cx_p;
                const R* cb_xy = C.Cb_ex_y + xg * cx_p; const R* cb_xz = C.Cb_ex_z + xg * cx_p;
                const R* ca_yz = C.Ca_ey_z + xg * cy_p; const R* ca_yx = C.Ca_ey_x + xg * cy_p;
                const R* cb_yz = C.Cb_ey_z + xg * cy_p; const R* cb_yx = C.Cb_ey_x + xg * cy_p;
                const R* ca_zx = C.Ca_ez_x + xg * cz_p; const R* ca_zy = C.Ca_ez_y + xg * cz_p;
                const R* cb_zx = C.Cb_ez_x + xg * cz_p; const R* cb_zy = C.Cb_ez_y + xg * cz_p;

                // ex_y, ex_z on rows 1..Ny-1, columns 1..Nz-1 of the (Ny+1, Nz+1) plane
                for (int j = 0; j < Nym1; ++j)
                    for (int k = 0; k < Nzm1; ++k) {
                        const int c = j * Nzm1 + k, e = (j + 1) * Nzp1 + (k + 1);
                        exy[e] = ca_xy[c] * exy[e] + cb_xy[c] * (hzp[(j + 1) * Nzp1 + (k + 1)] - hzp[j * Nzp1 + (k + 1)]);
                    }
                for (int j = 0; j < Nym1; ++j)
                    for (int k = 0; k < Nzm1; ++k) {
                        const int c = j * Nzm1 + k, e = (j + 1) * Nzp1 + (k + 1);
                        exz[e] = ca_xz[c] * exz[e] + cb_xz[c] * (hyp[(j + 1) * Nz + (k + 1)] - hyp[(j + 1) * Nz + k]);
                    }
                // ey_z, ey_x on rows 0..Ny-1, columns 1..Nz-1 of the (Ny, Nz+1) plane
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nzm1; ++k) {
                        const int c = j * Nzm1 + k, e = j * Nzp1 + (k + 1);
                        eyz[e] = ca_yz[c] * eyz[e] + cb_yz[c] * (hxp[j * Nz + k + 1] - hxp[j * Nz + k]);
                    }
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nzm1; ++k) {
                        const int c = j * Nzm1 + k, e = j * Nzp1 + (k + 1);
                        eyx[e] = ca_yx[c] * eyx[e] + cb_yx[c] * (hz1[e] - hzp[e]);
                    }
                // ez_x, ez_y on rows 1..Ny-1, columns 0..Nz-1 of the (Ny+1, Nz) plane
                for (int j = 0; j < Nym1; ++j)
                    for (int k = 0; k < Nz; ++k) {
                        const int c = j * Nz + k, e = (j + 1) * Nz + k;
                        ezx[e] = ca_zx[c] * ezx[e] + cb_zx[c] * (hy1[e] - hyp[e]);
                    }
                for (int j = 0; j < Nym1; ++j)
                    for (int k = 0; k < Nz; ++k) {
                        const int c = j * Nz + k, e = (j + 1) * Nz + k;
                        ezy[e] = ca_zy[c] * ezy[e] + cb_zy[c] * (hxp[(j + 1) * Nz + k] - hxp[j * Nz + k]);
                    }
                for (int e = 0; e < ex_p; ++e) ext[e] = exy[e] + exz[e];
                for (int e = 0; e < ey_p; ++e) eyt[e] = eyz[e] + eyx[e];
                for (int e = 0; e < ez_p; ++e) ezt[e] = ezx[e] + ezy[e];
            }
            // E soft sources and probes, solver.cpp:963-976 (every probe is treated as a source:
            // the reference sets is_source from a non-null dict lookup)
            for (size_t i = 0; i < n_e; ++i) {
                R* v = e_vals[i];
                if (e_probes[i]->is_source) {
                    *p_s1[i] = *p_s1[i] + v[n];
                    *p_s2[i] = *p_s2[i] + v[n];
                    *p_tot[i] = *p_s1[i] + *p_s2[i];
                }
                v[n] = *p_tot[i];
            }
        }
        {
            std::unique_lock<std::mutex> lk(mu);
            ++e_count;
            cv_ctrl.notify_all();
            cv_work.wait(lk, [&] { return e_done; });
        }
        if (ok) {
            // ---- H half step, solver.cpp:992-1101 ----
            for (int x = 0; x < nx; ++x) {
                R* hxy = hx_y + (size_t)x * hx_p; R* hxz = hx_z + (size_t)x * hx_p; R* hxt = hx + (size_t)x * hx_p;
                R* hyz = hy_z + (size_t)x * hy_p; R* hyx = hy_x + (size_t)x * hy_p; R* hyt = hy + (size_t)x * hy_p;
                R* hzx = hz_x + (size_t)x * hz_p; R* hzy = hz_y + (size_t)x * hz_p; R* hzt = hz + (size_t)x * hz_p;
                const R* exp_ = ex + (size_t)x * ex_p;
                const R* eyp = ey + (size_t)x * ey_p;
                const R* ezp = ez + (size_t)x * ez_p;
                const R* ey0 = (x > 0) ? ey + (size_t)(x - 1) * ey_p : halo[t - 1].ey;
                const R* ez0 = (x > 0) ? ez + (size_t)(x - 1) * ez_p : halo[t - 1].ez;
                const size_t xg = (size_t)(x_start + x);
                const R* da_xy = C.Da_hx_y + xg * dx_p; const R* da_xz = C.Da_hx_z + xg * dx_p;
                const R* db_xy1 = C.Db_hx_y1 + xg * dx_p; const R* db_xy2 = C.Db_hx_y2 + xg * dx_p;
                const R* db_xz1 = C.Db_hx_z1 + xg * dx_p; const R* db_xz2 = C.Db_hx_z2 + xg * dx_p;
                const R* da_yz = C.Da_hy_z + xg * dy_p; const R* da_yx = C.Da_hy_x + xg * dy_p;
                const R* db_yz1 = C.Db_hy_z1 + xg * dy_p; const R* db_yz2 = C.Db_hy_z2 + xg * dy_p;
                const R* db_yx1 = C.Db_hy_x1 + xg * dy_p; const R* db_yx2 = C.Db_hy_x2 + xg * dy_p;
                const R* da_zx = C.Da_hz_x + xg * dz_p; const R* da_zy = C.Da_hz_y + xg * dz_p;
                const R* db_zx1 = C.Db_hz_x1 + xg * dz_p; const R* db_zx2 = C.Db_hz_x2 + xg * dz_p;
                const R* db_zy1 = C.Db_hz_y1 + xg * dz_p; const R* db_zy2 = C.Db_hz_y2 + xg * dz_p;

                // every H value of the plane is updated; association is a + (b - c)
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nz; ++k) {
                        const int h = j * Nz + k;
                        hxy[h] = da_xy[h] * hxy[h] + (db_xy2[h] * ezp[(j + 1) * Nz + k] - db_xy1[h] * ezp[j * Nz + k]);
                    }
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nz; ++k) {
                        const int h = j * Nz + k;
                        hxz[h] = da_xz[h] * hxz[h] + (db_xz2[h] * eyp[j * Nzp1 + k + 1] - db_xz1[h] * eyp[j * Nzp1 + k]);
                    }
                for (int j = 0; j < Nyp1; ++j)
                    for (int k = 0; k < Nz; ++k) {
                        const int h = j * Nz + k;
                        hyz[h] = da_yz[h] * hyz[h] + (db_yz2[h] * exp_[j * Nzp1 + k + 1] - db_yz1[h] * exp_[j * Nzp1 + k]);
                    }
                for (int h = 0; h < hy_p; ++h)
                    hyx[h] = da_yx[h] * hyx[h] + (db_yx2[h] * ezp[h] - db_yx1[h] * ez0[h]);
                for (int h = 0; h < hz_p; ++h)
                    hzx[h] = da_zx[h] * hzx[h] + (db_zx2[h] * eyp[h] - db_zx1[h] * ey0[h]);
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nzp1; ++k) {
                        const int h = j * Nzp1 + k;
                        hzy[h] = da_zy[h] * hzy[h] + (db_zy2[h] * exp_[(j + 1) * Nzp1 + k] - db_zy1[h] * exp_[j * Nzp1 + k]);
                    }
                for (int h = 0; h < hx_p; ++h) hxt[h] = hxy[h] + hxz[h];
                for (int h = 0; h < hy_p; ++h) hyt[h] = hyz[h] + hyx[h];
                for (int h = 0; h < hz_p; ++h) hzt[h] = hzx[h] + hzy[h];
            }
            // H probes, solver.cpp:1104-1116
            for (size_t i = 0; i < h_probes.size(); ++i) {
                R* v = h_vals[i];
                const size_t q = n_e + i;
                if (h_probes[i]->is_source) {
                    *p_s1[q] = *p_s1[q] + v[n];
                    *p_s2[q] = *p_s2[q] + v[n];
                    *p_tot[q] = *p_s1[q] + *p_s2[q];
                }
                v[n] = *p_tot[q];
            }
        }
        {
            std::unique_lock<std::mutex> lk(mu);
            ++h_count;
            cv_ctrl.notify_all();
            cv_work.wait(lk, [&] { return h_done; });
        }
        if (!ok) continue;
        // ---- monitors, solver.cpp:1130-1161 and 1168-1225 ----
        R* base[6] = {ex, ey, ez, hx, hy, hz};
        for (const MonitorSlot& m : monitors) {
            if (n > 0 && (n % m.n_step) != 0) continue;
            if (m.axis == 0 && (m.position < x_start || m.position >= x_stop)) continue;
            const int mn = n / m.n_step;
            const int N1N2 = m.N1 * m.N2;
            const R* f = base[m.field];
            if (m.n_phasors == 0) {
                R* dst = (R*)m.values + (size_t)mn * N1N2;
                if (m.axis == 0) {
                    const R* src = f + (size_t)(m.position - x_start) * N1N2;
                    for (int e = 0; e < N1N2; ++e) dst[e] = src[e];
                } else {
                    for (int r = 0; r < nx; ++r)
                        for (int c = 0; c < m.N2; ++c)
                            dst[(size_t)(x_start + r) * m.N2 + c] =
                                f[m.yz_offset + (size_t)r * m.row_stride + (size_t)c * m.col_stride];
                }
            } else {
                // values[f] += field * phase[mn, f]; the field is promoted to a complex with zero
                // imaginary part, so each component is one product and one sum.
                const std::complex<R>* ph = (const std::complex<R>*)m.phase;
                for (int fq = 0; fq < m.n_phasors; ++fq) {
                    std::complex<R>* acc = (std::complex<R>*)m.values + (size_t)fq * N1N2;
                    const std::complex<R> w = ph[(size_t)mn * m.n_phasors + fq];
                    const R wr = w.real(), wi = w.imag();
                    auto add = [&](size_t e, R v) {
                        R re = acc[e].real() + v * wr;
                        R im = acc[e].imag() + v * wi;
                        acc[e] = std::complex<R>(re, im);
                    };
                    if (m.axis == 0) {
                        const R* src = f + (size_t)(m.position - x_start) * N1N2;
                        for (int e = 0; e < N1N2; ++e) add(e, src[e]);
                    } else {
                        for (int r = 0; r < nx; ++r)
                            for (int c = 0; c < m.N2; ++c)
                                add((size_t)(x_start + r) * m.N2 + c,
                                    f[m.yz_offset + (size_t)r * m.row_stride + (size_t)c * m.col_stride]);
                    }
                }
            }
        }
    }

    // export this slab into the global Yee-shaped arrays (not part of the reference, which leaves
    // the fields in mem): order ex_y ex_z ey_z ey_x ez_x ez_y hx_y hx_z hy_z hy_x hz_x hz_y ex ey ez hx hy hz
    if (ok && fields_out) {
        R* slab[18] = {ex_y, ex_z, ey_z, ey_x, ez_x, ez_y, hx_y, hx_z, hy_z, hy_x, hz_x, hz_y, ex, ey, ez, hx, hy, hz};
        const int comp[18] = {EX, EX, EY, EY, EZ, EZ, HX, HX, HY, HY, HZ, HZ, EX, EY, EZ, HX, HY, HZ};
        for (int a = 0; a < 18; ++a) {
            if (!fields_out[a]) continue;
            const int c = comp[a];
            const int plane = (c == EX) ? ex_p : (c == EY) ? ey_p : (c == EZ) ? ez_p : (c == HX) ? hx_p : (c == HY) ? hy_p : hz_p;
            const int shift = (c == EY || c == EZ || c == HX) ? 1 : 0;  // slab plane x holds global x+shift
            for (int x = 0; x < nx; ++x)
                std::memcpy(fields_out[a] + (size_t)(x_start + x + shift) * plane, slab[a] + (size_t)x * plane,
                            sizeof(R) * plane);
        }
    }
}

template <typename R>
void Oracle<R>::run(int n_threads, int update_interval) {
    // solver.cpp:534-597: slabs of Nx / n_threads planes, the remainder spread over the first slabs;
    // zero dummy planes stand in for the neighbours of the outermost slabs.
    halo.assign(n_threads + 2, Halo<R>{});
    halo[0].ez = take((uint64_t)(Ny + 1) * Nz);
    halo[0].ey = take((uint64_t)Ny * (Nz + 1));
    halo[n_threads + 1].hy = take((uint64_t)(Ny + 1) * Nz);
    halo[n_threads + 1].hz = take((uint64_t)Ny * (Nz + 1));
    std::memset(halo[0].ez, 0, sizeof(R) * (size_t)(Ny + 1) * Nz);
    std::memset(halo[0].ey, 0, sizeof(R) * (size_t)Ny * (Nz + 1));
    std::memset(halo[n_threads + 1].hy, 0, sizeof(R) * (size_t)(Ny + 1) * Nz);
    std::memset(halo[n_threads + 1].hz, 0, sizeof(R) * (size_t)Ny * (Nz + 1));

    std::thread ctrl(&Oracle<R>::controller, this, n_threads, update_interval);
    std::vector<std::thread> th;
    const int batch = Nx / n_threads;
    int rem = Nx % n_threads;
    int x1 = 0;
    for (int t = 0; t < n_threads; ++t) {
        const int x0 = x1;
        x1 = x0 + batch + (rem > 0 ? 1 : 0);
        if (rem > 0) --rem;
        th.emplace_back(&Oracle<R>::worker, this, x0, x1, t + 1);
    }
    for (auto& t : th) t.join();
    ctrl.join();
    if (!thread_errors.empty()) throw OracleError{thread_errors.front()};
}

template <typename R>
int run_oracle(R* const coeff[30], oracle_probe* probes, int n_probes, oracle_monitor* monitors, int n_monitors,
               R* mem, uint64_t mem_len, int Nx, int Ny, int Nz, int Nt, int n_threads, int update_interval,
               R* const fields_out[18]) {
    try {
        if (Nx < 2 || Ny < 2 || Nz < 2 || Nt < 1) throw OracleError{"Invalid grid or step count."};
        if (n_threads <= 0 || n_threads > kMaxThreads) throw OracleError{"Invalid number of threads."};
        if (n_threads > Nx) throw OracleError{"More threads than x planes."};
        if (n_monitors > kMaxMonitors) throw OracleError{"Exceeded maximum number of monitors."};
        if (n_probes > kMaxProbes) throw OracleError{"Exceeded maximum number of probes."};
        for (int i = 0; i < 30; ++i)
            if (!coeff[i]) throw OracleError{"Missing coefficient array."};
        if (!mem) throw OracleError{"Memory buffer not initialized."};

        Oracle<R> o;
        o.Nx = Nx; o.Ny = Ny; o.Nz = Nz; o.Nt = Nt;
        o.pool_next = mem;
        o.pool_left = mem_len;
        o.fields_out = fields_out;
        Coeffs<R>& C = o.C;
        R* const* c = coeff;
        C.Ca_ex_y = c[0]; C.Ca_ex_z = c[1]; C.Cb_ex_y = c[2]; C.Cb_ex_z = c[3];
        C.Ca_ey_z = c[4]; C.Ca_ey_x = c[5]; C.Cb_ey_z = c[6]; C.Cb_ey_x = c[7];
        C.Ca_ez_x = c[8]; C.Ca_ez_y = c[9]; C.Cb_ez_x = c[10]; C.Cb_ez_y = c[11];
        C.Da_hx_y = c[12]; C.Da_hx_z = c[13]; C.Db_hx_y1 = c[14]; C.Db_hx_y2 = c[15]; C.Db_hx_z1 = c[16]; C.Db_hx_z2 = c[17];
        C.Da_hy_z = c[18]; C.Da_hy_x = c[19]; C.Db_hy_z1 = c[20]; C.Db_hy_z2 = c[21]; C.Db_hy_x1 = c[22]; C.Db_hy_x2 = c[23];
        C.Da_hz_x = c[24]; C.Da_hz_y = c[25]; C.Db_hz_x1 = c[26]; C.Db_hz_x2 = c[27]; C.Db_hz_y1 = c[28]; C.Db_hz_y2 = c[29];

        // solver.cpp:294-328: zero the last x plane of Cy, Cz and Dx (PEC wall at x = Nx).
        const size_t cy_p = (size_t)Ny * (Nz - 1), cz_p = (size_t)(Ny - 1) * Nz, dx_p = (size_t)Ny * Nz;
        for (R* a : {C.Cb_ey_z, C.Cb_ey_x, C.Ca_ey_z, C.Ca_ey_x}) std::memset(a + (Nx - 1) * cy_p, 0, sizeof(R) * cy_p);
        for (R* a : {C.Cb_ez_x, C.Cb_ez_y, C.Ca_ez_x, C.Ca_ez_y}) std::memset(a + (Nx - 1) * cz_p, 0, sizeof(R) * cz_p);
        for (R* a : {C.Db_hx_y1, C.Db_hx_y2, C.Db_hx_z1, C.Db_hx_z2, C.Da_hx_y, C.Da_hx_z})
            std::memset(a + (Nx - 1) * dx_p, 0, sizeof(R) * dx_p);

        // solver.cpp:351-463, CPU layout (gpu = 0)
        const int f_Ny[6] = {Ny + 1, Ny, Ny + 1, Ny, Ny + 1, Ny};
        const int f_Nz[6] = {Nz + 1, Nz + 1, Nz, Nz, Nz, Nz + 1};
        for (int m = 0; m < n_monitors; ++m) {
            const oracle_monitor& pm = monitors[m];
            if (pm.field < 0 || pm.field > 5 || pm.axis < 0 || pm.axis > 2 || pm.n_step < 1)
                throw OracleError{"Invalid monitor."};
            MonitorSlot s{};
            s.field = pm.field; s.axis = pm.axis; s.position = pm.position; s.n_step = pm.n_step;
            s.n_phasors = pm.n_frequencies; s.values = pm.values; s.phase = pm.dtft_phase;
            if (!s.values || (s.n_phasors > 0 && !s.phase)) throw OracleError{"Invalid monitor arrays."};
            if (s.axis == 0) {
                if (s.field == HX || s.field == EY || s.field == EZ) s.position -= 1;
                s.row_stride = f_Nz[s.field]; s.col_stride = 1; s.yz_offset = 0;
                s.N1 = f_Ny[s.field]; s.N2 = f_Nz[s.field];
            } else if (s.axis == 1) {
                s.row_stride = f_Ny[s.field] * f_Nz[s.field]; s.col_stride = 1;
                s.yz_offset = s.position * f_Nz[s.field];
                s.N1 = Nx; s.N2 = f_Nz[s.field];
            } else {
                s.row_stride = f_Ny[s.field] * f_Nz[s.field]; s.col_stride = f_Nz[s.field];
                s.yz_offset = s.position;
                s.N1 = Nx; s.N2 = f_Ny[s.field];
            }
            o.monitors.push_back(s);
        }
        // solver.cpp:465-532
        for (int i = 0; i < n_probes; ++i) {
            const oracle_probe& pp = probes[i];
            if (pp.field < 0 || pp.field > 5 || !pp.values) throw OracleError{"Invalid probe."};
            ProbeSlot s{pp.field, pp.idx[0], pp.idx[1], pp.idx[2], pp.is_source != 0};
            if (s.field == EY || s.field == EZ || s.field == HX) s.x -= 1;
            if (s.field == EX || s.field == EZ || s.field == HY) s.y -= 1;
            if (s.field == EX || s.field == EY || s.field == HZ) s.z -= 1;
            if (s.x < 0 || s.y < 0 || s.z < 0) throw OracleError{"Source cannot be placed on grid edge"};
            o.probes.push_back(s);
            o.probe_values.push_back((R*)pp.values);
        }
        o.run(n_threads, update_interval);
        return 0;
    } catch (OracleError& e) {
        g_err = e.msg;
        return 1;
    } catch (std::exception& e) {
        g_err = e.what();
        return 2;
    }
}

}  // namespace

extern "C" {

int oracle_run_f32(float* const coefficients[30], oracle_probe* probes, int n_probes, oracle_monitor* monitors,
                   int n_monitors, float* mem, uint64_t mem_len, int Nx, int Ny, int Nz, int Nt, int n_threads,
                   int update_interval, float* const fields_out[18]) {
    return run_oracle<float>(coefficients, probes, n_probes, monitors, n_monitors, mem, mem_len, Nx, Ny, Nz, Nt,
                             n_threads, update_interval, fields_out);
}

int oracle_run_f64(double* const coefficients[30], oracle_probe* probes, int n_probes, oracle_monitor* monitors,
                   int n_monitors, double* mem, uint64_t mem_len, int Nx, int Ny, int Nz, int Nt, int n_threads,
                   int update_interval, double* const fields_out[18]) {
    return run_oracle<double>(coefficients, probes, n_probes, monitors, n_monitors, mem, mem_len, Nx, Ny, Nz, Nt,
                              n_threads, update_interval, fields_out);
}

}  // extern "C"
