/*
 * cntc_amd.h -- C ABI of the MI355X contact-solve library (libcntc_amd.so).
 *
 * Module-3 subset of CONTACT's cntc_* library interface: same symbol names (lower case, no
 * trailing underscore), same argument lists (every argument by pointer), same units and error
 * codes.  A caller of the reference library (the Fortran ISO_C_BINDING interface
 * src/contact_addon.ifc, the C prototypes matlab_intfc/contact_addon.h:7-148 or the ctypes
 * wrapper python_intfc/__init__.py) links this library instead.  Each declaration cites the
 * reference routine it replaces.
 *
 * Scope of the solver behind cntc_calculate: module 3 (basic contact), elastic half-space with
 * quasi-identical materials (K = 0), Coulomb friction (L = 0), uniform grids, creepages or
 * tangential forces prescribed (F = 0, 1, 2), T = 0..3 with chi = 0, normal problem by NormCG,
 * steady rolling by GDsteady (G=5, with the reference's SteadyGS fall-back) or SteadyGS (G=0,
 * 3, 4), shift and transient rolling by TangCG, and ConvexGS for G=2 and for steady rolling
 * with contact in the trailing-edge column.  Settings outside this subset make cntc_calculate
 * return CNTC_err_other (-99) instead of solving.
 */
#ifndef CNTC_AMD_H
#define CNTC_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- initialisation -------------------------------------------------------------------- */

/* src/contact_addon.f90:237 cntc_initializeFirst.  The output folder and experiment name of
   the first call name the .inp file written for CNTC_if_wrtinp (<outdir>/<expnam>.inp, expnam
   "contact_addon" when none is given). */
void cntc_initializefirst(int *ifcver, int *ierror, int *ioutput, const char *c_wrkdir,
                          const char *c_outdir, const char *c_expnam, int *len_wrkdir,
                          int *len_outdir, int *len_expnam);
/* src/contact_addon.f90:476-576 cntc_initialize (imodul must be 3); the output folder of the
   first initialising call is kept, as in the reference */
void cntc_initialize(int *ire, int *imodul, int *ifcver, int *ierror, const char *c_outdir,
                     int *len_outdir);
/* src/m_caddon_data.f90:310-332 cntc_getMaxNumThreads (1: calls are serialised; batches of
   problems go through cntc_calculatebatch) */
void cntc_getmaxnumthreads(int *mxthrd);
/* src/contact_addon.f90:580 cntc_setGlobalFlags */
void cntc_setglobalflags(int *lenflg, int *params, int *values);

/* ---- case set-up ----------------------------------------------------------------------- */

/* src/contact_addon.f90:780-973 cntc_setFlags */
void cntc_setflags(int *ire, int *icp, int *lenflg, int *params, int *values);
/* src/contact_addon.f90:1128-1283 cntc_setSolverFlags */
void cntc_setsolverflags(int *ire, int *icp, int *gdigit, int *nints, int *iparam, int *nreals,
                         double *rparam);
/* src/contact_addon.f90:1287-1473 cntc_setMaterialParameters (M=0) */
void cntc_setmaterialparameters(int *ire, int *icp, int *mdigit, int *nparam, double *rparam);
/* src/contact_addon.f90:1554 cntc_setTimestep */
void cntc_settimestep(int *ire, int *icp, double *dt);
/* src/contact_addon.f90:1587 cntc_setReferenceVelocity */
void cntc_setreferencevelocity(int *ire, int *icp, double *veloc);
/* src/contact_addon.f90:1619-1667 cntc_setRollingStepsize */
void cntc_setrollingstepsize(int *ire, int *icp, double *chi, double *dq);
/* src/contact_addon.f90:1671 cntc_setFrictionMethod (L=0) */
void cntc_setfrictionmethod(int *ire, int *icp, int *imeth, int *nparam, double *params);
/* src/contact_addon.f90:1972-2093 cntc_setHertzContact (ipotcn -1, -2, -3; -4, -5 and -6 are
   accepted with the reference's parameter counts but make cntc_calculate return
   CNTC_err_other: this library does not implement the prescribed-pressure options) */
void cntc_sethertzcontact(int *ire, int *icp, int *ipotcn, int *nparam, double *rparam);
/* src/contact_addon.f90:2097-2257 cntc_setPotContact (ipotcn 1..4) */
void cntc_setpotcontact(int *ire, int *icp, int *ipotcn, int *nparam, double *rparam);
/* src/contact_addon.f90:2294-2323 cntc_setPenetration */
void cntc_setpenetration(int *ire, int *icp, double *pen);
/* src/contact_addon.f90:2327-2357 cntc_setNormalForce */
void cntc_setnormalforce(int *ire, int *icp, double *fn);
/* src/contact_addon.f90:2361-2487 cntc_setUndeformedDistc (ibase 1, 9) */
void cntc_setundeformeddistc(int *ire, int *icp, int *ibase, int *nparam, double *prmudf);
/* src/contact_addon.f90:2597-2650 cntc_setTangentialForces: Fx, Fy relative to mu*Fn for the
   F digit (CNTC_ic_force) 1 or 2; |(Fx, Fy)| > 1 is scaled to 1 */
void cntc_settangentialforces(int *ire, int *icp, double *fx, double *fy);
/* src/contact_addon.f90:2491-2530 cntc_setCreepages */
void cntc_setcreepages(int *ire, int *icp, double *vx, double *vy, double *phi);
/* src/contact_addon.f90:2534-2593 cntc_setExtraRigidSlip: extra term (wx, wy) of the tangential
   right-hand side per element of the potential contact (sets the E digit to 1) */
void cntc_setextrarigidslip(int *ire, int *icp, int *lenarr, double *wx, double *wy);
/* src/contact_addon.f90:1055-1124 cntc_setMetadata: metadata codes CNTC_mt_*, kept with the
   problem; they do not change the solution */
void cntc_setmetadata(int *ire, int *icp, int *lenmta, int *params, double *values);
/* src/contact_addon.f90:4001-4160 cntc_getFlags: control digits of the problem */
void cntc_getflags(int *ire, int *icp, int *lenflg, int *params, int *values);
/* src/contact_addon.f90:4164-4258 cntc_getParameters: itask 1 (veloc, chi, dq, spinxo, spinyo,
   tau_c0), 2 (material), 3 (friction) */
void cntc_getparameters(int *ire, int *icp, int *itask, int *lenarr, double *values);
/* src/contact_addon.f90:4987-5017 cntc_getReferenceVelocity */
void cntc_getreferencevelocity(int *ire, int *icp, double *veloc);

/* ---- calculation ----------------------------------------------------------------------- */

/* src/contact_addon.f90:3519-3552, 3754-3900 cntc_calculate / cntc_calculate3 */
void cntc_calculate(int *ire, int *icp, int *ierror);

/* No counterpart in the reference (additive entry point): solve the cases of `nprob`
   (ire, icp) problems, each set up with the calls above, in one batched run on the GPU.
   Replaces the reference's pattern of one cntc_calculate per OpenMP thread
   (src/test_table.f90:196-292); ierror[k] is what cntc_calculate would return. */
void cntc_calculatebatch(int *nprob, const int *ire, const int *icp, int *ierror);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops).  Solves the current case of (ire, icp) as
   cntc_calculate does, allocating nothing on the device: every device buffer of the solve
   (the patch's grid functions, coefficients, spectra and FFT workspaces, twiddle tables, and
   the staging, partials, counters and results of the batched launches) is carved from one
   fixed array of 64 MiB per GPU, reserved when the library sets up (cntc_initialize must have
   been called).  Limit: patches whose buffers exceed 56 MiB (about 160 x 160 elements) return
   CNTC_err_other, as do calls while the serving loop runs.  The solution is kept in host memory
   with the problem, so getters and later cases (continuation, warm starts) work as after
   cntc_calculate.  Calls share the array: two calls overlapping on one GPU would race (the
   library serialises its calls within a process). */
void cntc_calculatenoalloc(int *ire, int *icp, int *ierror);

/* ---- continuous batching (no counterpart in the reference) ------------------------------ */

/* For a multibody caller that issues patch solves as they arise.  cntc_servestart starts the
   serving loop: queued requests are admitted between batched launches while others are in
   flight, up to *budget_bytes of device memory of active patches (<= 0: 200 GB), and each
   request completes as soon as its own solve ends.  cntc_submit queues the current case of
   (ire, icp) and returns at once with *ticket > 0 (or an error code < 0); it starts the loop
   if needed.  cntc_poll reports *done = 1 with the outcome, 0 while in flight, -1 for an
   unknown ticket; cntc_wait blocks.  *latency is submit-to-result in seconds.  Any set-call
   or getter on a problem in flight first waits for its result.  cntc_servestop finishes all
   queued requests and stops the loop; cntc_calculate and cntc_calculatebatch go through the
   loop while it runs. */
void cntc_servestart(double *budget_bytes, int *ierror);
void cntc_submit(int *ire, int *icp, int *ticket);
void cntc_poll(int *ticket, int *done, int *ierror, double *latency);
void cntc_wait(int *ticket, int *ierror, double *latency);
void cntc_servestop(void);

/* ---- results --------------------------------------------------------------------------- */

/* src/contact_addon.f90:5021-5066 cntc_getHertzContact */
void cntc_gethertzcontact(int *ire, int *icp, int *lenarr, double *values);
/* src/contact_addon.f90:5070 cntc_getNumElements */
void cntc_getnumelements(int *ire, int *icp, int *mx, int *my);
/* src/contact_addon.f90 cntc_getGridDiscretization */
void cntc_getgriddiscretization(int *ire, int *icp, double *dx, double *dy);
/* src/contact_addon.f90 cntc_getPotContact (ipotcn=1 form: mx, my, xc1, yc1, dx, dy) */
void cntc_getpotcontact(int *ire, int *icp, int *lenarr, double *values);
/* src/contact_addon.f90 cntc_getPenetration */
void cntc_getpenetration(int *ire, int *icp, double *pen);
/* src/contact_addon.f90 cntc_getCreepages */
void cntc_getcreepages(int *ire, int *icp, double *vx, double *vy, double *phi);
/* src/contact_addon.f90:5275-5318 cntc_getContactForces */
void cntc_getcontactforces(int *ire, int *icp, double *fn, double *tx, double *ty, double *mz);
/* src/contact_addon.f90:5919-6007 cntc_getSensitivities: sens[iout + lenout*iin] =
   d output / d input, outputs (Fn, Fx/(mu Fn), Fy/(mu Fn), Mz), inputs (pen, cksi, ceta, cphi),
   in the caller's units; computed when CNTC_ic_sens is 2 (dFn/dpen) or 3 (also the
   creepages), with the G=6 solver parameters of cntc_setsolverflags.  Entries beyond 4x4 are
   left unchanged. */
void cntc_getsensitivities(int *ire, int *icp, int *lenout, int *lenin, double *sens);
/* src/contact_addon.f90:5460-5497 cntc_getContactPatchAreas */
void cntc_getcontactpatchareas(int *ire, int *icp, double *carea, double *harea, double *sarea);
/* src/contact_addon.f90:5501-5552 cntc_getElementDivision */
void cntc_getelementdivision(int *ire, int *icp, int *lenarr, int *eldiv);
/* src/contact_addon.f90:5556 cntc_getMaximumPressure */
void cntc_getmaximumpressure(int *ire, int *icp, double *pnmax);
/* src/contact_addon.f90:5591 cntc_getMaximumTraction */
void cntc_getmaximumtraction(int *ire, int *icp, double *ptmax);
/* src/contact_addon.f90:5672+ cntc_getFieldData (h, mu, px, py, pn, ux, uy, un, taucrt,
   uplsx, uplsy, sx, sy, wx, wy; an unknown code leaves fld unchanged, as in the reference) */
void cntc_getfielddata(int *ire, int *icp, int *ifld, int *lenarr, double *fld);
/* src/contact_addon.f90:5889-5915 cntc_getDisplacements: displacement differences un, ux, uy of
   the solution (A ps on the interior, 0 elsewhere) */
void cntc_getdisplacements(int *ire, int *icp, int *lenarr, double *un, double *ux, double *uy);
/* src/contact_addon.f90:5827-5855 cntc_getTractions */
void cntc_gettractions(int *ire, int *icp, int *lenarr, double *pn, double *px, double *py);
/* src/contact_addon.f90 cntc_getMicroslip */
void cntc_getmicroslip(int *ire, int *icp, int *lenarr, double *sx, double *sy);
/* No counterpart in the reference, which prints these counts in its flow output
   (src/m_snorm.f90:329-331, src/m_stang.f90:57): values = {NORM outer iterations, NormCG
   iterations, TANG outer iterations, tangential solver iterations, solver flags} of the last
   case; flags bit 0: GDsteady stagnated, bit 1: SteadyGS ran, bit 2: solved among the largest
   patches of a batch, with flushes of their own. */
void cntc_getiterations(int *ire, int *icp, int *lenarr, int *values);
/* src/contact_addon.f90:6011-6040 cntc_getCalculationTime (for cntc_calculatebatch, twall is
   the time from the start of the batch to this problem's result) */
void cntc_getcalculationtime(int *ire, int *icp, double *tcpu, double *twall);
/* src/contact_addon.f90:6044-6067 cntc_resetCalculationTime */
void cntc_resetcalculationtime(int *ire, int *icp);

/* ---- kernel timing (no counterpart in the reference) ------------------------------------ */

/* For benchmark and profiling callers: mode 1 records HIP events around every kernel launch
   of subsequent calculations (resetting earlier records), mode 2 only around the SteadyGS,
   GD line-search probe and FFT product kernels, mode 0 stops recording.  Safe to call while
   the serving loop runs. */
void cntc_amdprofile(int *mode);
/* Per kernel (kinds[k] = operation*4 + stage): seconds, algorithmic bytes and launches
   accumulated since recording started; *nkinds receives the number of kernels recorded. */
void cntc_amdprofileget(int *lenarr, int *nkinds, int *kinds, double *seconds, double *bytes,
                        double *launches);

/* ---- finalisation ---------------------------------------------------------------------- */

/* src/contact_addon.f90:6191-6243 cntc_finalize */
void cntc_finalize(int *ire);
/* src/contact_addon.f90:6294 cntc_finalizeLast */
void cntc_finalizelast(void);

/* ---- flag codes (values of src/caddon_flags.inc:11-178) -------------------------------- */
#define CNTC_if_units   1933
#define CNTC_un_cntc    1934
#define CNTC_un_spck    1935
#define CNTC_un_si      1936
#define CNTC_ic_config  1967
#define CNTC_ic_pvtime  1970
#define CNTC_ic_bound   1971
#define CNTC_ic_tang    1972
#define CNTC_ic_norm    1973
#define CNTC_ic_force   1974
#define CNTC_ic_frclaw  1976
#define CNTC_ic_discns  1977
#define CNTC_ic_inflcf  1978
#define CNTC_ic_mater   1979
#define CNTC_ic_exrhs   1980
#define CNTC_ic_xflow   1981
#define CNTC_ic_heat    1982
#define CNTC_ic_iestim  1983
#define CNTC_ic_output  1984
#define CNTC_ic_flow    1985
#define CNTC_ic_return  1986
#define CNTC_ic_matfil  1987
#define CNTC_ic_sens    1988
#define CNTC_ic_ifmeth  1989
#define CNTC_ic_ifvari  1990
#define CNTC_ic_sbsout  1991
#define CNTC_ic_sbsfil  1992
#define CNTC_ic_npomax  1993
#define CNTC_if_idebug  2000
/* >= 1 per problem: cntc_calculate appends the case to the .inp file in the layout of the
   reference's wrtinp (src/m_sinput.f90:1506-1862), for replay by the stand-alone program */
#define CNTC_if_wrtinp  2002
#define CNTC_if_openmp  2003
#define CNTC_if_ncase   2005
/* No counterpart in the reference: 0 = FP64 products (default), 1 = FP32 FFT products inside
   the FP64 refinement (residual, projections and forces stay FP64). */
#define CNTC_if_precision 2009

#define CNTC_fld_h    1
#define CNTC_fld_mu   2
#define CNTC_fld_px   3
#define CNTC_fld_py   4
#define CNTC_fld_pn   5
#define CNTC_fld_ux   7
#define CNTC_fld_uy   8
#define CNTC_fld_un   9
#define CNTC_fld_taucrt 11
#define CNTC_fld_uplsx 12
#define CNTC_fld_uplsy 13
#define CNTC_fld_sx   15
#define CNTC_fld_sy   16
#define CNTC_fld_wx   22
#define CNTC_fld_wy   23

/* metadata codes of cntc_setmetadata (src/caddon_flags.inc:155-163) */
#define CNTC_mt_tim   2011
#define CNTC_mt_xr    2012
#define CNTC_mt_yr    2013
#define CNTC_mt_xw    2014
#define CNTC_mt_yw    2015
#define CNTC_mt_sw    2016
#define CNTC_mt_run   2021
#define CNTC_mt_axle  2022
#define CNTC_mt_side  2023

#define CNTC_err_allow  (-12)
#define CNTC_err_norm   (-27)
#define CNTC_err_tang   (-28)
#define CNTC_err_icp    (-31)
#define CNTC_err_other  (-99)

#ifdef __cplusplus
}
#endif
#endif
