// ops.hpp -- per-patch work items of the batched device operations.
//
// Every device operation of the engine is launched once for all patches that asked for it in
// the same scheduler round.  Each participating patch contributes one item; the work-groups of
// a launch are split between items by a prefix table (item i owns groups [gstart[i],
// gstart[i+1])).  Reductions are deterministic: every group writes its partial sums, and the
// item's last group to finish combines them in an order fixed by the item's group count.
#pragma once

#include <cstdint>

namespace cntc {

// Where a launch's reductions go: group partials [group * K + k]; with `count` set (one
// counter per item, zero between launches) the last group of each item also combines them
// into results [item * K + k] with the item's combine mode (modes[item]: 0 sum, 1 max, 2 min).
struct RedOut {
    double* partials;
    unsigned* count;
    double* results;
    const int* modes;
};

// element-state codes, src/m_gridfunc.f90:30-37
enum : int8_t { EXTER = 0, ADHES = 1, SLIP = 2, PLAST = 3 };

// element selections, src/m_gridfunc.f90:25-33 (AllElm, AllInt) plus single states
enum Mask : int { M_ALL = 0, M_INT = 1, M_EXT = 2, M_ADH = 3, M_SLP = 4 };

enum OpKind : int {
    OP_VEC = 0,
    OP_NCG_NEGPN,
    OP_NCG_DD,
    OP_NCG_EXPAND,
    OP_NCG_HMIN,
    OP_SN_CONTRACT,
    OP_SN_EXPAND,
    OP_SN_FINAL,
    OP_AIJ_ONES,
    OP_TCG_NT,
    OP_TCG_RSET,
    OP_TCG_ZPREP,
    OP_TCG_QPOST,
    OP_TCG_PUPD,
    OP_TCG_AS,
    OP_TCG_SA,
    OP_ST_AS,
    OP_ST_SA,
    OP_GD_DP,
    OP_GD_PROJECT,
    OP_GD_PROBE,
    OP_GD_UPDATE,
    OP_GD_MOVES,
    OP_GD_DIAGRES,
    OP_CONV,
    OP_CONV32,
    OP_ELASCF,
    OP_EMBED,
    OP_EXTRACT,
    OP_SETUP,
    OP_FORCES,
    OP_STDYGS,
    OP_GUARD,
    OP_SENS,
    OP_SENS_EXT,
    OP_LEDGE,
    OP_GSB_SETUP,
    OP_GSB_WALK,
    OP_GSB_CHECK,
    OP_CNVXGS,
    OP_DENSE,
    OP_TCG_VUPD,
    OP_COUNT_KINDS
};

// sub-operations of OP_VEC (masked BLAS-1 over one or two components)
enum VecSub : int {
    V_SET = 0,      // x = a
    V_COPY,         // y = x
    V_SCAL,         // x *= a
    V_AXPY,         // y += a x
    V_AXPBY,        // y = a x + b y
    V_ADDC,         // x += a
    V_DOT,          // sum x*y, count
    V_SUM,          // sum x, count
    V_SUMSQ,        // sum x*x, count
    V_MAXABS,       // max |x|
    V_MIN,          // min x
    V_MAX,          // max x
    V_COUNT,        // count of selected elements
    V_NSUB
};

struct VecItem {
    int n, sub, mask, ncomp;
    const int8_t* el;
    double* x[2];
    double* y[2];
    double a, b;
    // device scalars (the patch's scalar area): sout receives the results (value, count) for
    // later launches; with sfn set, a (and b) are formed on the device from sin as the host
    // would form them (VecScalarFn; flag carries a host-side condition); echo[0..1] are
    // returned as results 2 and 3 of a sum, so that the host can form the same values
    double* sout;
    const double* sin[3];
    const double* echo[2];
    int sfn, flag;
};

// scalar functions of VecItem.sfn (NormCG, src/m_solvpn.f90:130-250)
enum VecScalarFn : int {
    SF_NONE = 0,
    SF_NEG_MEAN,    // a = -sin[0][0] / max(1, sin[0][1])  (minus the mean of a sum)
    SF_NCG_BETA,    // rz2 = sin[0][0], rzr = sin[1][0], rz1 = sin[2][0]: y = x if rz1 < TINY, else
                    // y = x + max(0, (rz2 - rzr) / max(TINY, rz1)) y
    SF_NCG_ALPHA,   // rv = sin[0][0], vav = sin[1][0]: a = rv / vav if |vav| > 1e-32 and flag,
                    // else rv / max(TINY, vav)
};
constexpr double NCG_TINY = 1e-20;

// NormCG / snorm element steps (src/m_solvpn.f90:298-393, src/m_snorm.f90:195-325)
struct NcgItem {
    int n, mode;     // k_ncg_hmin: 1 = enter where h < hmin (p kept), 2 = h <= hmin (p = 0)
    int8_t* el;
    double* p;       // pn
    double* dd;      // A ps - rhs
    const double* rhs;
    double* res;
    double* v;
    const double* h; // hstot
    const double* unn;
    double avg;      // mean subtracted from dd on the interior (proj_avg)
    double davg;     // penetration estimate (N=1) or 0
    double errpn, pen, tol, hmin;
};

// sum over the row extents of one element's influence row, with unit tractions
// (gf3_AijPj with p = 1, src/m_aijpj.f90:158-185)
struct AijItem {
    int mx, my, ix, iy;   // 1-based element
    int cf_mx, cf_my, ik, jk;
    const int8_t* el;
    const double* cf;     // coefficient block (ik,jk), layout [(iy+cf_my)*2cf_mx + ix+cf_mx]
    double ga_inv;
};

// TangCG steps (src/m_solvpt.f90:1841-2442)
struct TcgItem {
    int n, pad;
    int8_t* el;
    double* px; double* py; const double* pn;
    double* sx; double* sy;
    double* nx; double* ny;
    double* rx; double* ry;
    double* zx; double* zy;
    double* vx; double* vy;
    double* qx; double* qy;
    double* poldx; double* poldy;
    const double* p0x; const double* p0y;   // sensitivity solve: stopping test on p - p0
    double mu, ga, c11, c22;
    double alpha;
    // device scalars of the iteration (the patch's scalar area): zprep stores (z.q, v.q) in
    // sc[0..1], qpost (r.v, v.q) in sc[2..3]; vupd forms beta and pupd alpha from them, as the
    // host would, so that one iteration needs no host round trip before pupd's results
    double* sc;
};

// stang outer-loop moves (src/m_stang.f90:433-508)
struct StItem {
    int n, pad;
    int8_t* el;
    double* px; double* py; const double* pn;
    const double* sx; const double* sy;
    double mu, tol, tol1, tol2, errpt;
};

// GDsteady (src/gdsteady.f90); row kernels use n = my rows, element kernels n = mx*my
struct GdItem {
    int n, mx, my, imeth;
    int kdown, first;
    int8_t* el;
    double* px; double* py; const double* pn;
    double* dpx; double* dpy;
    double* sx; double* sy;
    const double* wsx; const double* wsy;
    double* nx; double* ny;
    double* rx; double* ry;
    double* dvx; double* dvy;
    double* vx; double* vy;
    double* qx; double* qy;
    double* dscl;
    double* poldx; double* poldy;
    double mu, alpha, fdecay, fac_v;
    double c11, c12, c21, c22;
    double d_ifc, d_lin, d_cns, d_slp, pow_s;
};


// sensitivity solves (sens_norm, src/m_solvpn.f90:531-650; sens_tang, src/m_solvpt.f90:544-796)
//   OP_SENS, mode 0: copy a patch state (element division and up to 6 grid functions)
//   OP_SENS, mode 1: ws = wb plus the change of the rigid slip for creepage `which`
//                    (0 cksi, 1 ceta, 2 cphi) by d, on the interior
//   OP_SENS_EXT:     extents of the contact area, returns (-min ix, max ix, -min iy, max iy)
struct SensItem {
    int n, mx, mode, nvec;
    int which, pad;
    const int8_t* el;          // mode 1, extents: element division
    const int8_t* el_src; int8_t* el_dst;
    const double* src[6]; double* dst[6];
    const double* wbx; const double* wby;
    double* wsx; double* wsy;
    double xc1, yc1, dx, dy, dq, d;
    const double* facdt;       // mode 1: fraction of the step in contact (null: 1)
};

// leading edge of transient rolling with dq > dx (sxbnd, stang_rhs (1b), src/m_leadedge.f90,
// src/m_stang.f90:838-935).  mode 0, one thread per row: facdt and the leading-edge position
// lix (0-based ix0, or -1 where facdt >= 0.9999).  mode 1, per element: the fixed right-hand
// side ws = -facdt h - (A' pv)(cv) on the interior, with the displacement at the leading edge
// interpolated from b = A pv (cs) between ix0 and ix0 + 1 where lix >= 0.
struct LedgeItem {
    int n, mx, my, mode;
    int steady, pad;           // mode 1, steady rolling: ws = -facdt h only
    const int8_t* el;
    double* facdt; double* lix;
    const double* hx; const double* hy;
    const double* ux; const double* uy;     // A' pv_t (cv), interior
    const double* bx; const double* by;     // A pv_t (cs), all elements
    double* wsx; double* wsy;
    double xc1, dx, dq, fxdfac;
};

// one FFT product (gf3_VecAijPj via fft_VecAijPj, src/m_aijpj.f90:258-451, 712-1015)
struct ConvItem {
    int mx, my, Lx, Ly;        // grid and padded real sizes; complex row length M = Lx/2
    int nin, nout, mask, mode; // mode 0: product, 1: forward transform only, 2: reciprocal
    int rpg, cpg;              // rows per work-group of the row stages, columns per work-group
                               // of the column stage (0: the default shapes); any shape gives
                               // the same bits, each transform's arithmetic is its own
    const double* in[3];
    double* out[3];
    const double* addend[3];   // added to the product where written (may be null)
    const void* spec[3][3];    // [out][in] spectra, layout [kx*Ly + ky], null = zero block
    const int8_t* el;
    void* wsS;                 // row spectra [c][y*H + kx]
    void* wsT;
    const void* twM;           // exp(-2 pi i e/M), e < M
    const void* twLx;          // exp(-2 pi i e/Lx), e < Lx
    const void* twLy;          // exp(-2 pi i e/Ly), e < Ly
    double scale;              // mode 1/2: factor applied to the spectrum
};

// SteadyGS (src/m_solvpt.f90:2825-3254): one work-group runs every sweep of one patch
struct GsItem {
    int n, mx, my, k;          // k = number of interior elements
    int maxgs, pad;
    double eps, omegah, omegas, ga_inv, mu, facnel;
    int8_t* el;
    double* px; double* py; const double* pn;
    double* sx; double* sy;
    const double* wsx; const double* wsy;
    double* dpx; double* dpy;  // work: differences along x
    int* list;                 // work: 2n ints, (element, jy*2mx + jx) of the summation range
    const double* cf[2][2];    // tangential blocks of cs, layout [(iy+my)*2mx + ix+mx]
    const double* p0x; const double* p0y;   // sensitivity solve: stopping test on p - p0
    // block variant (GsbItem): the FFT product z = A dp and the block change d (work vectors),
    // and the product's item (spectra of cs, workspaces, twiddles; in, out, addend set per use)
    double* zx; double* zy;
    double* ddx; double* ddy;
    ConvItem conv;
};

// SteadyGS by blocks of rows (the sweeps and element updates of OP_STDYGS, src/m_solvpt.f90:
// 2825-3254).  The influence of the rows outside the block being walked comes from the FFT
// product z = A dp (cs), formed at each sweep's start and updated by the change d of each
// walked block; only rows inside a block are summed directly.  One work-group per patch.
//   mode 0 (set-up): dp = differences along x, the row ranges, d = 0, the sweep state;
//     returns (first interior row, last interior row)
//   mode 1 (walk): rows [y0, y1) in order; first clears d on rows [z0, z1)
//   mode 2 (check, after a sweep): the convergence test; returns
//     (itgs, dif, flops of the direct sums, 0, 0, dif of the first sweep, still sweeping, 0..)
struct GsbItem {
    int n, mx, my, k;
    int mode, maxgs;
    int y0, y1, z0, z1;
    double eps, omegah, omegas, ga_inv, mu, facnel;
    int8_t* el;
    double* px; double* py; const double* pn;
    double* sx; double* sy;
    const double* wsx; const double* wsy;
    double* dpx; double* dpy;
    double* zx; double* zy;
    double* ddx; double* ddy;
    int* rng;                  // 4 my ints: summation range [rj0, rj1], interior [ri0, ri1]
    const double* cf[2][2];
    const double* p0x; const double* p0y;
    double* st;                // 4 doubles: squared changes of the sweep, sweeps, dif1, flops
};

// ConvexGS (src/m_solvpt.f90:2446-2821, chi = 0, no plasticity): one wavefront per patch runs
// every sweep.  Point Gauss-Seidel over the interior elements in grid order, each shift by
// direct summation with the current tractions: s = ws + A(cs) p (T=1, 2), s = ws + A(cs - cv) p
// in steady rolling, and near its leading edge (lix >= 0) s = ws + A(cs) p - ubnd with ubnd
// from A(cs) p at the sweep's start.  Returns (itgs, dif, flops, 0, 0, dif1, 0, 0, 0, 0).
struct CgsItem {
    int n, mx, my, k;
    int maxgs, steady;
    double eps, omegah, omegas, ga_inv, mu, facnel;
    int8_t* el;
    double* px; double* py; const double* pn;
    double* sx; double* sy;
    const double* wsx; const double* wsy;
    const double* lix;         // steady rolling: 0-based leading-edge position, or -1
    double* ubx; double* uby;  // work: ubnd per element
    int* list;                 // work: the interior elements (k)
    const double* cf[2][2];    // tangential blocks of cs, layout [(iy+my)*2mx + ix+mx]
    const double* cv[2][2];    // the same blocks of cv (steady rolling)
    const double* p0x; const double* p0y;   // sensitivity solve: stopping test on p - p0
};

// Dense influence product of a small patch (n <= DENSE_MAX_N) on FP64 MFMA: the same product
// as OP_CONV mode 0, out_i = addend_i + (1/G) sum_j A_ij in_j over the (out, in) blocks, written
// where the mask holds (src/m_aijpj.f90:99-254, 424-448: the reference's direct summation).
// The block-Toeplitz matrix is applied as a sum over the row offsets d = iy - jy of ordinary
// matrix products (the Toeplitz block of offset d times the source rows jy = iy - d).
constexpr int DENSE_MAX_N = 2500;
struct DenseItem {
    int mx, my, nin, nout, mask, pad;
    const double* in[3];
    double* out[3];
    const double* addend[3];   // added to the product where written (may be null)
    const double* cf[3][3];    // [out][in] coefficient tables [(iy+my)*2mx + ix+mx], null = zero
    const int8_t* el;
    double ga_inv;
};

// per-case fields from the case parameters (set_prev_data, set_norm_rhs, set_tang_rhs and
// init_curr_data, src/m_sdis.f90:208-768); returns (max h, max -h).  Forces: returns
// (sum pn, sum px, sum py, sum px*y, sum py*x).
struct SetupItem {
    int n, mx, my, ibase;          // ibase 1: quadratic gap from prm; otherwise h is given
    int prev, zero_pn, zero_pt, hertz;   // hertz: 1 elliptical (3D), 2 strip |y| <= bb (2D)
    int exrhs, regularize;         // exrhs: hx, hy hold the extra rigid slip on entry;
                                   // regularize: I=1 (exterior 0, slip on the bound)
    double mu;
    double facx, facy;             // spin factors of the rigid slip (E digit 2, 3 zero one)
    int8_t* el;
    double* h; double* hx; double* hy;
    double* pn; double* px; double* py;
    double* pvn; double* pvx; double* pvy;
    double xc1, yc1, dx, dy;
    double prm[6];
    double cksi, ceta, cphi, dq, xofs, yofs;
    double hz_aa, hz_bb, pnmax;
};

// generation of the influence coefficients (elascf_pcwcns, src/m_visc.f90:431-604)
struct CoefItem {
    int mx, my, pad0, pad1;
    double dx, dy, xshft, yshft, akv, nuv;
    double* cf;   // 9 blocks of (2mx)(2my)
};

// embedding of a coefficient block into the padded real array of a transform
struct EmbedItem {
    int Lx, Ly, rx, ry;        // padded sizes; offsets d kept for -rx <= d < rx (and y)
    int cf_mx, cf_my, pad0, pad1;
    const double* cf;          // block in [(iy+cf_my)*2cf_mx + ix+cf_mx] layout
    double* out;               // Lx*Ly real, x fastest
    double scale;
};

// extraction of a circulant result back into a coefficient block
struct ExtractItem {
    int Lx, Ly, cf_mx, cf_my;
    const double* in;          // Lx*Ly
    double* cf;                // block
    double scale;
};

}  // namespace cntc
