// engine.hpp -- host side of the HIP contact solver: case inputs, per-patch device state and
// the batched scheduler that runs many patch solves through shared kernel launches.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace cntc {

// Inputs of one case, in CONTACT's internal units (mm, N, N/mm^2).  Defaults are the
// module-3 defaults of the cntc_* library (src/m_hierarch_data.f90:966-1014, 1728-1747,
// src/m_global_data.f90:336-350).
struct CaseInput {
    int tang = 3, norm = 0, pvtime = 2, iestim = 0, gausei = 0;
    // B digit: 0 normal problem solved; 2 Hertzian elliptical pressure prescribed, normal
    // problem skipped (src/m_sdis.f90:641-709, src/m_scontc.f90:409-411)
    int bound = 0;
    double poiss1 = 0.28, poiss2 = 0.28, gg1 = 1.0, gg2 = 1.0;
    double fstat = 0.30;
    int ipotcn = 0;                 // 1..4 grid forms, -1/-2/-3 Hertzian; 0 = not yet set
    // set when the last potential-contact call asked for an option this library does not solve
    // (Hertzian rectangles and SDEC, ipotcn -4..-6): calculate reports it instead of solving
    // the previous geometry
    const char* unsupported = nullptr;
    int mx = 1, my = 1;
    double xl = 0, yl = 0, xh = 0, yh = 0, xc1 = 0, yc1 = 0, xcm = 0, ycm = 0, dx = 1, dy = 1;
    double hz_a1 = 0, hz_b1 = 0, hz_aa = 0, hz_bb = 0, hz_aob = 1.0, hz_scale = 1.0, hz_bneg = 0, hz_bpos = 0;
    int ibase = 0;                  // 1 quadratic, 9 element-wise; 0 = not yet set
    double prmudf[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> h_elem;
    double pen = 0, fn = 0, cksi = 0, ceta = 0, cphi = 0, chi = 0, dq = -1.0, veloc = 1e3;
    int maxgs = 999, maxin = 20, maxnr = 25, maxout = 1;
    double eps = 1e-5;
    double fdecay = (double)0.90f, betath = (double)0.10f, d_ifc = 1.0, d_lin = 1.0, d_cns = 1.0,
           d_slp = 1.0, pow_s = (double)0.90f;
    int kdowfb = 1;
    double omegah = (double)0.90f, omegas = 1.0;   // SteadyGS relaxation for G=2, 3
    int precision = 0;              // 0: FP64 products; 1: FP32 products in FP64 refinement
    // sensitivities: S2 digit (2: dFn/dpen, 3: also the creepages) and their solver
    // parameters (src/m_hierarch_data.f90:981, 1733-1735)
    int sens = 1, mxsens = 30;
    double epsens = 1e-4;
    // F digit: 0 creepages prescribed; 1 Fx, 2 Fx and Fy prescribed (fractions of mu Fn), the
    // creepages cksi (and ceta) then being solved for and kept in cksi, ceta
    int force3 = 0;
    double fxrel = 0, fyrel = 0;
    // E digit (src/m_sdis.f90:517-573): 1 adds the extra rigid slip exrhs (x then y, mx*my each,
    // cntc_setextrarigidslip); 2 drops the spin term from wy, 3 from wx
    int rztang = 0;
    std::vector<double> exrhs;
};

// Results of one case.
struct CaseResult {
    int ierror = 0;
    double fn = 0, fx = 0, fy = 0, mz = 0;
    double pen = 0;
    int itnorm = 0, ittang = 0, itcg = 0, itgs = 0;
    int gd_stagnated = 0;
    int used_stdygs = 0;
    int prio = 0;                    // solved with flushes of its own (largest patches of a batch)
    double t_done = 0;               // seconds from the start of the solve call to this result
    double hz_a1 = 0, hz_b1 = 0, hz_aa = 0, hz_bb = 0, hz_rho = 0, hz_cp = 0;
};

struct DeviceArena {
    char* base = nullptr;
    size_t bytes = 0;
    ~DeviceArena();
    void ensure(size_t n);
};

// One contact problem (a result element / contact patch pair of the library): inputs of the
// last case, the grid, and the device-resident grid functions that persist between cases.
struct Patch {
    CaseInput in;
    CaseResult out;
    int mx = 0, my = 0, n = 0;
    double dx = 1, dy = 1, dxdy = 1, xc1 = 0, yc1 = 0;
    int Lx = 0, Ly = 0;
    bool have_solution = false;
    bool cnvx = false;               // the tangential solver of the case in progress is ConvexGS
    bool dense = false;              // influence products by the dense FP64-MFMA kernel (n <= 2500)
    // What the coefficient tables and spectra in the arena (cf, cfv, spec, specv, spec32 and the
    // preconditioner's cfm, specm, specm32) were computed for.  A case with the same key reuses
    // them; prec_valid has bit kk set when preconditioner block kk is there too.
    struct CoefKey {
        const void* cf = nullptr;
        int mx = 0, my = 0, Lx = 0, Ly = 0, cvmode = 0, precision = 0;
        double dx = 0, dy = 0, ak = 0, nu = 0, ga = 0, chi = 0, dq = 0;
        bool operator==(const CoefKey& o) const
        {
            return cf == o.cf && mx == o.mx && my == o.my && Lx == o.Lx && Ly == o.Ly &&
                   cvmode == o.cvmode && precision == o.precision && dx == o.dx && dy == o.dy &&
                   ak == o.ak && nu == o.nu && ga == o.ga && chi == o.chi && dq == o.dq;
        }
    };
    CoefKey coef_key;
    bool coef_valid = false;
    unsigned prec_valid = 0;
    int slot = -1;                   // scheduler slot during a solve
    // sensitivities sens(iout, iin) in internal units, iout + 4 iin: outputs fn, fx/(mu fn),
    // fy/(mu fn), mz; inputs pen, cksi, ceta, cphi (src/m_hierarch_data.f90:122-124)
    double sens[16] = {};
    // material and kinematics of the current case
    double ga = 1, nu = 0, ak = 0, muscal = 0.3, fntrue = 0, pen = 0, dq = 1, chi = 0;
    bool nt_cpl = false;
    std::vector<double> h_host;      // undeformed distance, host copy where needed
    std::vector<int8_t> el_host;
    double hmin = 0, hmax = 0;
    int ibase_eff = 1;
    double prm[6] = {0, 0, 0, 0, 0, 0};
    bool hertz_init = false;
    double hz_pnmax = 0, xofs = 0, yofs = 0;
    // device memory
    DeviceArena mem;
    size_t mem_need = 0;
    int8_t* el = nullptr;
    double *pn = nullptr, *px = nullptr, *py = nullptr, *sx = nullptr, *sy = nullptr;
    double *pvx = nullptr, *pvy = nullptr, *pvn = nullptr;
    double *h = nullptr, *hx = nullptr, *hy = nullptr, *wsx = nullptr, *wsy = nullptr;
    double* w[24] = {};              // work vectors
    double* hn = nullptr;            // the normal gap NormCG solves for: h, plus A_nt p_t when K != 0
    double* po[3] = {};              // tractions of the previous outer (Panagiotopoulos) iteration
    double *un = nullptr, *ux = nullptr, *uy = nullptr;   // displacement differences A ps
    double *facdt = nullptr, *lix = nullptr;  // leading edge of transient rolling (P.ledge)
    bool ledge = false;              // T=2 with dq > dx
    double *cf = nullptr, *cfv = nullptr, *cfm = nullptr;  // coefficient blocks cs, cv, ms
    void *spec = nullptr, *specv = nullptr, *specm = nullptr, *spec32 = nullptr, *specm32 = nullptr;
    double* embed = nullptr;         // padded real scratch for spectra (Lx*Ly, or 4*cf_mx*cf_my)
    void *wsS = nullptr, *wsT = nullptr;
    double* sc = nullptr;            // device scalars that one launch leaves for the next (64)
    const void *twM = nullptr, *twLx = nullptr, *twLy = nullptr;
    const void *twM32 = nullptr, *twLx32 = nullptr, *twLy32 = nullptr;
    const void *ptwM = nullptr, *ptwLx = nullptr, *ptwLy = nullptr;   // preconditioner sizes
    // After a solve that allocated nothing (Engine::solve_noalloc) the solution lives here, as
    // the leading block of the arena (element states, then the 40 grid functions) laid out as
    // in device memory; the next solve that uses the arena loads it back.
    std::vector<char> host_state;
    bool state_on_host = false;
};

class Scheduler;

// a request of the continuous-batching loop
struct Ticket {
    Patch* p = nullptr;
    int state = 0;                   // 0 queued, 1 admitted, 2 done
    int rc = 0;
    double t_submit = 0, t_done = 0;
};

// device bytes of one patch's arena
size_t patch_bytes(int mx, int my);

// The engine owns the stream, the staging buffers and the twiddle-table cache.
class Engine {
public:
    Engine();
    ~Engine();
    // Solve one case for every patch (each with its own inputs).  Returns 0 when every case
    // ran; individual outcomes are in Patch::out.
    int solve(const std::vector<Patch*>& patches);
    // Solve one case allocating no device memory: every device buffer of the solve (the
    // patch's arena, twiddle tables, staging, partials, counters, results) is carved from a
    // fixed array of noalloc_bytes() per GPU, set up with the engine.  Returns 0 when the case
    // ran (outcome in Patch::out), -1 when it cannot run this way (last_error says why: the
    // patch does not fit, the serving loop is running).
    int solve_noalloc(Patch* p);
    static size_t noalloc_bytes();
    static size_t noalloc_patch_limit();   // bytes of arena and twiddle tables that fit
    hipStream_t stream() const { return stream_; }
    bool noalloc_active() const { return noalloc_; }
    // download the element data of a patch
    void get_field(const Patch& p, int what, double* out) const;
    void get_eldiv(const Patch& p, int* out) const;
    // twiddle tables exp(-2 pi i e / L), e < L, cached per length and precision
    const void* twiddles(int L, bool fp32);
    // tile shapes of the FFT products of one size class (CNTC_FFT_TUNE=0: the defaults), timed
    // once per class on synthetic data of that size: {rows per work-group, columns per
    // work-group}, 0 for a default
    std::pair<int, int> conv_shape(bool fp32, int Lx, int Ly, int nin, int nout);
    std::string last_error;
    std::chrono::steady_clock::time_point t_solve0;   // start of the current solve call
    // per-kernel timing (HIP events around every launch); key = kind*4 + stage
    void set_profile(int mode);
    int get_profile(int len, int* kinds, double* seconds, double* bytes, long* launches) const;
    long rounds() const;

    // Continuous batching.  A server thread owns the stream and the scheduler while it runs:
    // it admits queued requests between flushes (up to `budget_bytes` of active patch
    // arenas, at least one), runs all admitted solves through the same batched launches and
    // completes each request as soon as its own solve ends.  solve() must not be called
    // while serving.
    bool serve_start(size_t budget_bytes);
    void serve_stop();
    bool serving() const { return serving_; }
    int submit(Patch* p);                                   // ticket > 0, -1 when not serving
    int poll(int ticket, int* rc, double* latency, bool wait);   // 1 done, 0 pending, -1 unknown
    void forget(int ticket);                                // drop a completed ticket

private:
    bool prepare(Patch& p);
    char* noalloc_base_ = nullptr;   // device address of the fixed array on this engine's GPU
    bool noalloc_ = false;           // prepare carves the arena from the fixed array
    void serve_loop();
    Patch* tickets_at(int id);
    std::thread server_;
    std::mutex smtx_;
    std::condition_variable scv_, dcv_;
    std::deque<int> queue_;
    std::map<int, Ticket> tickets_;
    int next_ticket_ = 1;
    bool serving_ = false, stop_ = false;
    size_t budget_ = 0, in_use_ = 0;
    hipStream_t stream_ = nullptr;
    std::unique_ptr<Scheduler> sched_;
    std::vector<std::pair<long, void*>> tw_;
    std::map<long, std::pair<int, int>> shapes_;   // conv_shape by (fp32, Lx, Ly, nin, nout)
};

int opt_fft_size(int n);

}  // namespace cntc
