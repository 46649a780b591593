/*
 * cntc_oracle.h -- CPU restatement (FP64, scalar C) of CONTACT's inner contact solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * contact_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  The product path never links, imports or calls it.
 *
 * It follows the reference's algorithm for one contact patch on a uniform grid:
 *   coefficients  src/m_visc.f90:127-604   (sgencr, elascf_pcwcns)
 *   convolution   src/m_aijpj.f90:99-1119  (gf3_AijPj, gf3_VecAijPj, fft_makePrec, opt_fft_size)
 *   normal        src/m_snorm.f90:31-378, src/m_solvpn.f90:24-527
 *   tangential    src/m_stang.f90:28-951, src/m_solvpt.f90:51-493, 1841-2442, 3258-3807,
 *                 src/gdsteady.f90:9-1462
 *   set-up        src/m_sdis.f90:125-830, src/m_hertz.f90:267-499, src/m_leadedge.f90:92-332
 *   outputs       src/m_soutpt.f90:424-450, src/contact_addon.f90:5275-5855
 * Function-level citations are in cntc_oracle.c.
 */
#ifndef CNTC_ORACLE_H
#define CNTC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* element-state codes, src/m_gridfunc.f90:30-37 */
enum { ORC_EXTER = 0, ORC_ADHES = 1, ORC_SLIP = 2, ORC_PLAST = 3 };

typedef struct orc_input {
    /* control digits (src/m_hierarch_data.f90:966-1014 and the module-3 defaults of
       src/m_global_data.f90:336-350) */
    int tang;      /* T: 0 frictionless, 1 shift, 2 transient rolling, 3 steady rolling */
    int norm;      /* N: 0 penetration prescribed, 1 normal force prescribed */
    int pvtime;    /* P: 0 continuation, 2 initiation */
    int iestim;    /* I: 0 zero initial estimate, 3 previous solution */
    int gausei;    /* G digit (5: GDsteady parameters in effect) */
    /* material, M=0: poisson ratios and shear moduli of both bodies [-], [N/mm2] */
    double poiss1, poiss2, gg1, gg2;
    /* Coulomb friction, L=0 */
    double fstat;
    /* potential contact: ipotcn=3 form (centre of first element and step sizes), or
       Hertzian ipotcn=-1 (curvatures a1, b1) / -2 (curvature a1, ellipticity aob) / -3
       (semi-axes aa, bb) / -4 (curvature a1, half-width bb) / -5 (half-sizes aa, bb), with scale */
    int ipotcn;
    int mx, my;
    double xc1, yc1, dx, dy;
    double hz_a1, hz_b1, hz_aa, hz_bb, hz_scale, hz_aob;
    /* undeformed distance: ibase=1 quadratic (prmudf[0..5]) or ibase=9 element-wise */
    int ibase;
    double prmudf[6];
    const double *h_elem;    /* ibase=9: mx*my values, x fastest */
    /* kinematics */
    double pen, fn, cksi, ceta, cphi, chi, dq, veloc;
    /* solver parameters, src/contact_addon.f90:1128-1283 */
    int maxgs, maxin, maxnr, maxout;
    double eps;
    double fdecay, betath, d_ifc, d_lin, d_cns, d_slp, pow_s;
    int kdowfb;
    /* relaxation of SteadyGS for G=3 (G=0, 4, 5 use the reference's defaults) */
    double omegah, omegas;
    /* sensitivities: S2 digit (CNTC_ic_sens, src/contact_addon.f90:844-846) and the G=6
       solver parameters (src/contact_addon.f90:1266-1272) */
    int sens, mxsens;
    double epsens;
    /* F digit (CNTC_ic_force, src/contact_addon.f90:834-842): 0 creepages prescribed, 1 Fx,
       2 Fx and Fy prescribed as fractions of mu*Fn (cntc_setTangentialForces, :2597-2650) */
    int force3;
    double fxrel, fyrel;
    /* E digit (src/m_sdis.f90:517-553): 1 adds exrhs (2*mx*my values, x then y, from
       cntc_setExtraRigidSlip, src/contact_addon.f90:2534-2593); 2 drops the spin term from wy,
       3 from wx */
    int rztang;
    const double *exrhs;
    /* B digit (src/m_sdis.f90:641-709, src/m_scontc.f90:409-411): 0 normal problem solved, 2
       Hertzian elliptical pressure prescribed (set by the Hertz options -4 and -5 as well) */
    int bound;
} orc_input;

/* fills an input record with the reference's module-3 defaults */
void orc_default_input(orc_input *inp);

typedef struct orc_problem orc_problem;

orc_problem *orc_new(void);
void orc_free(orc_problem *p);

/* solve one case; the problem keeps its state (tractions, element division) for continuation
   (P=0, I=3).  Returns the cntc_calculate error code (0, CNTC_err_norm=-27, CNTC_err_tang=-28,
   or -99 for inputs the oracle does not cover). */
int orc_calculate(orc_problem *p, const orc_input *inp);

/* results */
int  orc_get_mx(const orc_problem *p);
int  orc_get_my(const orc_problem *p);
void orc_get_grid(const orc_problem *p, double *xc1, double *yc1, double *dx, double *dy);
void orc_get_forces(const orc_problem *p, double *fn, double *fx, double *fy, double *mz);
void orc_get_tractions(const orc_problem *p, double *pn, double *px, double *py);
void orc_get_slip(const orc_problem *p, double *sx, double *sy);
void orc_get_eldiv(const orc_problem *p, int *el);
void orc_get_h(const orc_problem *p, double *h);
double orc_get_pen(const orc_problem *p);
/* displacement differences un, ux, uy of the last solution (A ps on the interior) */
void orc_get_displacements(orc_problem *p, double *un, double *ux, double *uy);
void orc_get_hertz(const orc_problem *p, double *a1, double *b1, double *aa, double *bb,
                   double *rho, double *cp);
void orc_get_iters(const orc_problem *p, int *itnorm, int *ittang, int *itcg, int *itgs);
/* bit 0: the last GDsteady call ended in stagnation (src/gdsteady.f90:560-585);
   bit 1: SteadyGS ran in the last case */
int  orc_get_stagnated(const orc_problem *p);
/* sensitivity matrix in internal units, sens[iout + 4*iin] with outputs (fn, fx/(mu fn),
   fy/(mu fn), mz) and inputs (pen, cksi, ceta, cphi), src/m_hierarch_data.f90:122-124 */
void orc_get_sens(const orc_problem *p, double *sens);
/* creepages of the last case: with F=1, 2 the Newton iteration's final cksi (and ceta) */
void orc_get_creepages(const orc_problem *p, double *cksi, double *ceta, double *cphi);

/* stand-alone building blocks, exposed for known-answer tests */
int  orc_opt_fft_size(int n);
/* coefficients cf[(ik*3+jk)][iy+my][ix+mx] for ix in [-mx,mx), iy in [-my,my); unscaled by 1/G */
void orc_elascf_pcwcns(double akv, double nuv, int mx, int my, double dx, double dy,
                       double xshft, double yshft, double *cf);
/* linear convolution u(i) = sum_j cf(i-j) p(j) on an mx*my grid, for a single block:
   method 0 = direct summation, 1 = zero-padded FFT at 2*opt_fft_size */
void orc_conv_block(int mx, int my, const double *cfblock, const double *p, double *u,
                    int method);
/* preconditioner block m = G^2 * IFFT(1/FFT(c)) on the 2mx*2my circulant, src/m_aijpj.f90:457-708 */
void orc_make_prec_block(int mx, int my, double ga, const double *cfblock, double *mblock);
/* complete elliptic integrals B(m), D(m) with mc = 1-m (Carlson duplication) */
void orc_ellip_bd(double mc, double *b, double *d);
/* Kalker's linear-theory creep coefficients C11, C22, C23 at a/b and poisson ratio nu,
   src/m_hertz.f90:1147-1247 */
void orc_linrol(double nu, double aob, double *cxx, double *cyy, double *cyz);
/* 1-D complex DFT of any length (forward sign -1), for checks against numpy */
void orc_fft(int n, const double *re_in, const double *im_in, double *re_out, double *im_out,
             int sign);

#ifdef __cplusplus
}
#endif
#endif
