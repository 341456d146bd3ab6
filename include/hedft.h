/*
 * hedft.h — C-ABI of the MI355X-native BCN-4He-DFT time-step hot path.
 *
 * The reference (s-kesh/He_DFT_2024) has no FFI; its "API" is a set of Fortran
 * subroutines that communicate through module arrays.  Every entry point below
 * replaces one of those subroutines (cited file:line, paths relative to the
 * reference root) and keeps its argument meaning.  Module arrays live on the
 * device inside a context; hedft_upload/hedft_download copy them to and from
 * the host in the reference's own layout (Fortran column-major, x fastest:
 * idx = ix + nx*(iy + ny*iz); complex(8) = interleaved re,im; half spectra are
 * (nx/2+1, ny, nz)), so the Fortran wrapper passes c_loc() of its arrays.
 *
 * Conventions: every function returns int (HEDFT_OK = 0 on success, a negative
 * HEDFT_E* code otherwise; hedft_strerror() names it).  No exceptions cross the
 * ABI.  A context owns one HIP stream; calls on one context are not
 * thread-safe, distinct contexts are independent.  Functions that return host
 * scalars synchronise the context's stream; all others only enqueue work.
 */
#ifndef HEDFT_H
#define HEDFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HEDFT_OK 0
#define HEDFT_EINVAL -1     /* bad argument (size, id, mode)                 */
#define HEDFT_EGRID -2      /* grid edge not supported (tstgrid.f90 values)  */
#define HEDFT_EHIP -3       /* HIP runtime error                             */
#define HEDFT_ENOMEM -4     /* device allocation failed                      */
#define HEDFT_ESTATE -5     /* call out of order (e.g. PC before RK start)   */
#define HEDFT_EBOUNDS -6    /* impurity left the interpolation table          */

/* ---- modes ------------------------------------------------------------- */
#define HEDFT_MODE_STATICS 0 /* real psi, imaginary time (He_statics)         */
#define HEDFT_MODE_TDDFT 1   /* complex psi, real time (Dynamics_TDDFT)       */

#define HEDFT_PREC_FP64 0    /* plain FP64 path (the drop-in)                 */
#define HEDFT_PREC_MIXED 1   /* FP32 mean field/stencils, FP64 psi + sums     */

#define HEDFT_MAX_IMP 4
#define HEDFT_MAX_QIMP 4 /* quantum impurities (wave functions) per context */

/* ---- field identifiers (module arrays of the reference) ------------------ */
enum hedft_field {
  HEDFT_DEN = 0,      /* rho:            real  N   (module rho)              */
  HEDFT_PSI,          /* psi:   real N (statics) / complex N (TDDFT)         */
  HEDFT_POT4,         /* field: pot4     real  N                             */
  HEDFT_DENCG,        /* rho: dencg      real  N                             */
  HEDFT_DELJ4,        /* lenard4: delj4  real  N                             */
  HEDFT_DENALF,       /* alphasterm      real  N                             */
  HEDFT_FALFS,        /*                 real  N                             */
  HEDFT_INTXALF,      /*                 real  N                             */
  HEDFT_INTYALF,      /*                 real  N                             */
  HEDFT_INTZALF,      /*                 real  N                             */
  HEDFT_UALPHAS,      /*                 real  N                             */
  HEDFT_DXDEN,        /* deriva          real  N                             */
  HEDFT_DYDEN,        /*                 real  N                             */
  HEDFT_DZDEN,        /*                 real  N                             */
  HEDFT_UIMP,         /* classicimp uimp real  N (evaluated on download)     */
  HEDFT_UEXT,         /* field: uext     real  N (statics, lexternal)        */
  HEDFT_TIMEC,        /* statics: real N (local step); TDDFT: complex N      */
  HEDFT_PSIOLD1,      /* rho: psiold(:,:,:,k) in PHYSICAL slot k = 1..3      */
  HEDFT_PSIOLD2,
  HEDFT_PSIOLD3,
  HEDFT_HPSIOLD1,     /* rho: hpsiold(:,:,:,k), k = 1..2                     */
  HEDFT_HPSIOLD2,
  HEDFT_PC,           /* rkpc: pc                                            */
  HEDFT_Q,            /* rkpc: q                                             */
  HEDFT_FDEN,         /* rho: fden   complex (nx/2+1,ny,nz)                  */
  HEDFT_FVLJ4,        /* lenard4: fvlj4 real (nx/2+1,ny,nz)                  */
  HEDFT_WCGK,         /* rho: wcgk      real (nx/2+1,ny,nz)                  */
  HEDFT_KALFS,        /* alphasterm: kalfs real (nx/2+1,ny,nz)               */
  HEDFT_WORK1,        /* scratch real N fields for the derivnD / FFT hooks   */
  HEDFT_WORK2,
  HEDFT_WORKC1,       /* scratch complex N field                             */
  HEDFT_WORKK1,       /* scratch complex half spectrum                       */
  /* quantum impurity (Dynamics_quantum_impurities, module quantal_imp); these
   * exist once hedft_set_quantum_impurity has been called                  */
  HEDFT_PSIX,         /* psix(:,:,:,1)   complex N                           */
  HEDFT_POTX4,        /* potx4           real  N                             */
  HEDFT_UPOTX,        /* upotx(:,:,:,1)  real  N                             */
  HEDFT_DENX,         /* denx = |psix|^2 real  N (evaluated on download)     */
  HEDFT_UIMPK,        /* uimp_k(:,:,:,1) real (nx/2+1,ny,nz)                 */
  HEDFT_PSIXOLD1,     /* psixold(:,:,:,k,1) in PHYSICAL slot k = 1..3        */
  HEDFT_PSIXOLD2,
  HEDFT_PSIXOLD3,
  HEDFT_HPSIXOLD1,    /* hpsixold(:,:,:,k,1), k = 1..2                       */
  HEDFT_HPSIXOLD2,
  HEDFT_PCX,          /* pcx(:,:,:,1)                                        */
  HEDFT_QX,           /* qx(:,:,:,1)                                         */
  HEDFT_PSIL,         /* statics: psil = log|psi| of the log-psi steppers    */
  HEDFT_NFIELDS
};

/* Grid: x_i = -xmax + h*(i-1), h = 2*xmax/nx (He_statics/BCN4HeDFT.f90:293-327). */
typedef struct hedft_grid {
  int32_t nx, ny, nz;
  int32_t pad_;
  double xmax, ymax, zmax;
} hedft_grid;

/* Functional and method parameters.  Defaults are the reference's module
 * initialisers (He_statics/modules.f90:15-19,109-133,214-226,297-299;
 * Dynamics_TDDFT/modules.f90 he4, DFT4He3d.f90:136-138); hedft_default_params
 * fills them. */
typedef struct hedft_params {
  double h2o2m4, cp4, cpp4, alphas, l, den0s;   /* he4                      */
  double csolid, beta_solid, den_m;             /* he4: C, beta, den_m      */
  double h4, eps4, sigma4;                      /* lenard4 (h4 = h4ot)      */
  int32_t core_otc;  /* 1: core4 = 'OTC' (alpha_s term), 0: 'OT '            */
  int32_t lsolid;    /* solid penalty; the drivers force core4='OT ' with it */
  int32_t npd;       /* stencil points: 13 (deriva: npd)                     */
  int32_t icon;      /* boundary mode of derivnD: 0, 8, 13, 18, 20, 24      */
  double dtmin, sdmu, dmumax, psimin;           /* statics local time step  */
  double lambdah, tmean[3], tsurf[3];           /* TDDFT absorbing layer    */
  int32_t precision; /* HEDFT_PREC_FP64 | HEDFT_PREC_MIXED                  */
  int32_t lexternal; /* statics: add uext to pot4 (He_statics/poten.f90:73-78) */
} hedft_params;

/* Per-term energies, He_statics/energy.f90:12-224 and
 * Dynamics_TDDFT/energy.f90:12-182 (module energies). */
typedef struct hedft_energies {
  double ekin4, elj4, ecor4, ealphas, esolid, eext; /* helium terms          */
  double etot4;                                     /* helium total          */
  double ekinx, eHeX, eimpu_impu, eimpu;            /* impurity terms        */
  double etot;                                      /* etot4 + eimpu         */
  double norm;                                      /* sum(den)*dxyz         */
} hedft_energies;

/* Droplet observables printed at the energy points of the TDDFT driver
 * (Dynamics_TDDFT/DFT4He3d.f90:709-836, r_cm.f90:1-30). */
typedef struct hedft_observables {
  double norm;     /* auxn4 = sum(den)*dxyz                                  */
  double rcm[3];   /* r_cm(den, n4): sum(den*r)*dxyz/n4, n4 the nominal N    */
  double l[3];     /* <Lx,Ly,Lz> about rcm: real part of sum(i conj(psi)     */
                   /* ((r-rcm) x grad) psi)*dxyz, first-derivative stencils  */
  double vcom[3];  /* real part of sum(-i conj(psi) grad psi)*dxyz/(norm*mhe) */
  double r2[3];    /* sum(den*x^2)*dxyz, y, z (about the origin)             */
} hedft_observables;

typedef struct hedft_ctx hedft_ctx;

const char* hedft_strerror(int rc);
/* Fills the reference defaults (OT functional, npd=13, icon=13, core 'OT ',
 * lsolid=.false., Dtmin=1e-10, Sdmu=0.1, Dmumax=100, Lambdah=2, t*mean=1000,
 * t*surf=2, FP64). */
void hedft_default_params(hedft_params* p);
/* 1 when n is an accepted grid edge (tstgrid.f90:1-23, the FFT sizes this
 * library plans), else 0. */
int hedft_grid_supported(int32_t n);

/* dimen.f90 + fftini (fft.f90:5-91) + fforma/initcg/kalfs kernel set-up
 * (fforma.f90:11-61, initcg.f90:32-49, BCN4HeDFT.f90:849).  Allocates every
 * module array of the chosen mode on device `device`. */
int hedft_create(const hedft_grid* g, const hedft_params* p, int32_t mode,
                 int32_t device, hedft_ctx** out);
int hedft_destroy(hedft_ctx* c);
int hedft_upload(hedft_ctx* c, int32_t field, const void* host, size_t bytes);
int hedft_download(hedft_ctx* c, int32_t field, void* host, size_t bytes);
size_t hedft_field_bytes(const hedft_ctx* c, int32_t field);
int hedft_synchronize(hedft_ctx* c);

/* ---- L1: numerical kernels (parity hooks) ----------------------------- */
/* derivnD(id,nn,h,iv,f,df,icon), init_deriv_parallel.f90:373-648 (real) and
 * :933-1209 (complex): id-th derivative along axis iv (1..3) of field `in`
 * into field `out`, with the context's icon and the grid step of that axis. */
int hedft_derivnD(hedft_ctx* c, int32_t id, int32_t iv, int32_t in, int32_t out);
/* fftfw(a,b) / fftfw_den: unnormalised R2C, e^{-i} (He_statics/fft.f90:37-59,
 * Dynamics_TDDFT/fft.f90:96-114).  in: real N field, out: spectrum field. */
int hedft_fftfw(hedft_ctx* c, int32_t in, int32_t out);
/* fftbk(c,d) / fftbk_*: C2R scaled by 1/(nx*ny*nz) (He_statics/fft.f90:62-89,
 * Dynamics_TDDFT/fft.f90:118-168).  The input spectrum is destroyed. */
int hedft_fftbk(hedft_ctx* c, int32_t in, int32_t out);

/* ---- L2: physics operators ------------------------------------------- */
/* derden(): grad(den) -> dxden,dyden,dzden (He_statics/derden.f90:5-25). */
int hedft_derden(hedft_ctx* c);
/* poten(): statics semantics (He_statics/poten.f90:5-221) use the stored
 * fden/dencg/denalf/grad(den); TDDFT semantics (Dynamics_TDDFT/poten.f90:124-221)
 * recompute them from den.  Includes term_alfa when core_otc. */
int hedft_poten(hedft_ctx* c);
/* term_alfa() alone (He_statics/term_alfa.f90:5-102); needs dencg/denalf and
 * grad(den) current.  Writes falfs, int*alf, ualphas. */
int hedft_term_alfa(hedft_ctx* c);
/* energy(): per-term energies (He_statics/energy.f90:12-224,
 * Dynamics_TDDFT/energy.f90:12-182). Synchronises. */
int hedft_energy(hedft_ctx* c, hedft_energies* e);
/* The observables block of the TDDFT driver (Dynamics_TDDFT/DFT4He3d.f90:727-836):
 * centre of mass (r_cm with the nominal atom number n4), angular momentum about
 * it, centre-of-mass velocity (mhe = 4.002572164*mp_u in K^-1 A^-2 in the
 * reference, DFT4He3d.f90:103,315) and <x^2>, <y^2>, <z^2>.  TDDFT contexts.
 * Synchronises. */
int hedft_compute_observables(hedft_ctx* c, int32_t n4, double mhe, hedft_observables* o);

/* ---- L3: integrators -------------------------------------------------- */
/* STEPRKR(deltat,mu4,mu4err,n4), He_statics/steprkr.f90:1-96. */
int hedft_steprkr(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4);
/* STEPPCR(deltat,mu4,mu4err,n4), He_statics/steppcr.f90:1-108; errpc is the
 * module variable ErrPC (may be NULL). */
int hedft_steppcr(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4,
                  double* errpc);
/* Log-psi variants of the imaginary-time steppers (Lrkpcl): psil = log|psi|
 * evolves as dpsil/dtau = -(H_l - mu), H_l = -(lap psil + |grad psil|^2) h2o2m4
 * + pot4, with one-sided stencils (Icon_local = 0, modules.f90:294), deltat
 * without the local time step, psil >= underpsil = -700 (modules.f90:298), and
 * psi = exp(psil) renormalised to n4 after every stage.
 * hedft_statics_log_start: psil = log|psi| (BCN4HeDFT.f90:1081-1083).
 * STEPRKRL(deltat,mu4,mu4err,n4), He_statics/steprkrl.f90:1-98.
 * STEPPCRL(deltat,mu4,mu4err,n4), He_statics/steppcrl.f90:1-116 (errpc: ErrPC).
 * hedft_statics_iteration_log: STEPRKRL when iterrk <= 4 else STEPPCRL, then the
 * driver refresh and poten (BCN4HeDFT.f90:1122-1160). */
int hedft_statics_log_start(hedft_ctx* c);
int hedft_steprkrl(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4);
int hedft_steppcrl(hedft_ctx* c, double deltat, double* mu4, double* mu4err, double n4, double* errpc);
int hedft_statics_iteration_log(hedft_ctx* c, double deltat, int32_t iterrk, double n4, double* mu4,
                                double* mu4err, double* errpc);
/* STEPRK(deltat), Dynamics_TDDFT/steprk.f90:1-144. */
int hedft_steprk(hedft_ctx* c, double deltat);
/* STEPPC(deltat,errHe,errimp,errvimp), Dynamics_TDDFT/steppc.f90:1-180.
 * Any of the error pointers may be NULL; when all are NULL nothing syncs. */
int hedft_steppc(hedft_ctx* c, double deltat, double* errHe, double* errimp, double* errvimp);

/* ---- classical impurities (Dynamics_TDDFT/potenimp*.f90) -------------- */
/* potenimpini (potenimpini.f90:31-59): radial tables potion(k,1:npot) with
 * r_i = (i-1)*delinter, rows contiguous per impurity; mass in K^-1 A^-2
 * (m_imp = m_imp_u*mp_u, DFT4He3d.f90:298). */
int hedft_set_impurities(hedft_ctx* c, int32_t nimp, int32_t npot, double delinter,
                         double rmaxinterpol, const double* potion, const double* mass);
/* Impurity-impurity terms of forceimp and energy (potenimp.f90:40-58,
 * energy.f90:168-177) for nimp >= 2, replacing the reference's
 * Select_pot(drselec_gs_k_k(k,m), r, ...) (force) and
 * Select_pot(selec_gs_k_k(k,m), r, ...) (energy) by radial tables on
 * r_i = i*delpair, i = 0..npair-1, interpolated as updatepoten interpolates
 * potion (potenimp.f90:101-112).  One table of npair values per pair k < m, in
 * the order (1,2),(1,3),..,(1,N),(2,3),..: vpair feeds eimpu_impu, dvpair the
 * pair force F_ij(k,m,:) = -dv(|R_k-R_m|) (R_k-R_m)/|R_k-R_m|, F_ij(m,k) =
 * -F_ij(k,m), added to F(k,:).  A distance beyond the last node is clamped and
 * sets lstopimp.  With nimp >= 2 the integrators, forceimp and the drivers
 * return HEDFT_ESTATE until the tables are set; hedft_set_impurities clears
 * them. */
int hedft_set_impurity_pairs(hedft_ctx* c, int32_t npair, double delpair, const double* vpair,
                             const double* dvpair);
/* rimp, vimp, aimp as (nimp,3) Fortran arrays (k fastest); NULL skips one. */
int hedft_set_impurity_state(hedft_ctx* c, const double* rimp, const double* vimp,
                             const double* aimp);
int hedft_get_impurity_state(hedft_ctx* c, double* rimp, double* vimp, double* aimp,
                             double* force, int32_t* lstopimp);
/* potenimp()/updatepoten (potenimp.f90:1-19,83-123): uimp follows rimp; the
 * device evaluates it on the fly from the tables, so this only validates. */
int hedft_potenimp(hedft_ctx* c);
/* forceimp() (potenimp.f90:23-67): F = -sum(grad(den)*uimp_k)*dxyz. */
int hedft_forceimp(hedft_ctx* c);
/* aimp = F/m (DFT4He3d.f90:692-696). */
int hedft_update_aimp(hedft_ctx* c);
/* forceimp() without a context and without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  Same
 * arithmetic as hedft_forceimp on a context holding the same fields (results
 * are bit-identical).  Every pointer except g is DEVICE memory of `device`:
 * dxden/dyden/dzden (nx*ny*nz FP64, x fastest), potion (nimp x npot tables, as
 * hedft_set_impurities), rimp (nimp,3, k fastest), dvpair (the pair force
 * tables of hedft_set_impurity_pairs, needed when nimp >= 2), force (out,
 * (nimp,3)), lstop (int flag set when a pair distance leaves its table; may be
 * NULL).  Work is enqueued on `stream` (a hipStream_t; NULL = the default
 * stream) and not waited for.  Scratch: a fixed per-GPU device array of
 * 1024 x 12 + 32 doubles held until the process ends; nimp above
 * HEDFT_MAX_IMP is rejected.  Calls that overlap on one GPU share that array
 * and race: serialise them (one stream, or events). */
int hedft_forceimp_noalloc(const hedft_grid* g, int32_t device, void* stream, const double* dxden,
                           const double* dyden, const double* dzden, int32_t nimp, int32_t npot,
                           double delinter, double rmaxinterpol, const double* potion,
                           const double* rimp, int32_t npair, double delpair, const double* dvpair,
                           double* force, int32_t* lstop);

/* ---- quantum impurity (Dynamics_quantum_impurities) -------------------- */
/* On a statics context this is the He_statics impurity (limp): uimp_k is vq
 * (BCN4HeDFT.f90:746-760), psix is real (upload HEDFT_PSIX before
 * hedft_statics_init), poten adds upotx = F^-1[fden vq] and potx4 =
 * F^-1[fdenx vq] (He_statics/poten.f90:94-107), energy adds ekinx and eimpu
 * (energy.f90:76-149), mu4 is unused and lexternal is rejected.
 * On a TDDFT context: one impurity described by a wave function psix (N_imp = 1, so no
 * impurity-impurity kernel uimp_k_k).  uimp_k: the X-He interaction's Fourier
 * transform on the half spectrum (nx/2+1,ny,nz), real, as potenimpini.f90:19-29
 * fills uimp_k(:,:,:,1) (bytes = 8*(nx/2+1)*ny*nz); m_imp_u: impurity mass in
 * u (h2o2mimp = h2o2m4*4.002603/m_imp_u, steprk.f90:86); mu4: the chemical
 * potential subtracted from pot4 in the helium equation (DFT4He3d.f90:183-189,
 * 0 unless Lbulk).  Contexts without classical impurities only.
 * Allocates psix, its PC history, potx4 and upotx (zero) and switches the
 * context's TDDFT entry points to the quantum-impurity equations:
 *   hedft_potenimp:  potx4, upotx (potenimp.f90:1-43)
 *   hedft_poten:     poten with pot4 += potx4 (poten.f90:5-155; fden from den)
 *   hedft_steprk/steppc: helium and psix together (steprk.f90:53-124,
 *                    steppc.f90:36-157); steppc's errimp is errpsix, errvimp 0
 *   hedft_energy:    ekinx, eHeX, eimpu = ekinx + eHeX (energy.f90:102-132)
 *   hedft_tddft_init / iteration / run_pc: qx = 0 (DFT4He3d.f90:617), then
 *                    potenimp + poten after each step (DFT4He3d.f90:667-677)
 * Upload psix (HEDFT_PSIX) before hedft_tddft_init. */
int hedft_set_quantum_impurity(hedft_ctx* c, const double* uimp_k, size_t bytes, double m_imp_u,
                               double mu4);
/* N_imp = nq quantum impurities on a TDDFT context (nq <= HEDFT_MAX_QIMP; a
 * statics context takes nq = 1 only), replacing potenimpini.f90:19-44 and the
 * N_imp loops of Dynamics_quantum_impurities: uimp_k holds nq half spectra
 * (uimp_k(:,:,:,i), bytes each), uimp_kk nq*nq half spectra laid out as the
 * reference's uimp_k_k(nx/2+1,ny,nz,N_imp,N_imp) (column-major: spectrum
 * (i, j), the potential impurity j exerts on impurity i, at offset
 * (i + j*nq) * bytes), and pair[i + j*nq] != 0 marks the pairs whose
 * Selec_gs_k_k(i,j) is not 'null' (i != j; uimp_kk and pair may be NULL: no
 * pair coupling); m_imp_u[i] in u.  Each
 * impurity has its own psix and PC history:
 *   potx4   = sum_i F^-1[F(|psix_i|^2) uimp_k_i]
 *   upotx_i = F^-1[fden uimp_k_i] + sum_{j != i} F^-1[F(|psix_j|^2) uimp_k_k(i,j)]
 * every term transformed back on its own and summed in real space in that
 * order (potenimp.f90:13-39); the steppers advance every psix with the same
 * slot rotation; steppc's errimp is sum_i c9 sum|pcx_i| / nxyz
 * (steppc.f90:115-139); hedft_energy's ekinx sums the impurities, eHeX uses
 * the spectral sum and eimpu_impu = sum_{k<m} sum(F^-1[F(|psix_k|^2)
 * uimp_k_k(k,m)] |psix_m|^2) dxyz (energy.f90:102-132).  The PSIX*, UPOTX and
 * UIMPK fields and hedft_quantum_centroid refer to the impurity selected by
 * hedft_select_quantum_impurity (0 after this call).  Upload every psix
 * before hedft_tddft_init. */
int hedft_set_quantum_impurities(hedft_ctx* c, int32_t nq, const double* uimp_k, const double* uimp_kk,
                                 const int32_t* pair, size_t bytes, const double* m_imp_u, double mu4);
/* The quantum impurity (0-based) the PSIX*, UPOTX, UIMPK fields and
 * hedft_quantum_centroid refer to (no counterpart in the reference, whose
 * arrays carry the impurity as a last index). */
int hedft_select_quantum_impurity(hedft_ctx* c, int32_t k);
/* Norm and centre of the impurity density, sum(|psix|^2)*dxyz and
 * sum(|psix|^2 r)*dxyz/norm (DFT4He3d.f90:682-700, the rimp_k file rows).
 * Synchronises. */
int hedft_quantum_centroid(hedft_ctx* c, double* norm, double* rcm);

/* ---- driver glue (order of DFT4He3d.f90:677-696, BCN4HeDFT.f90:1113-1160) */
/* Statics start (BCN4HeDFT.f90:724-874): normalise den to n4, psi=sqrt(den),
 * fden, dencg, denalf, derden, term_alfa, poten; *mu4 = <psi|H|psi>/n4. */
int hedft_statics_init(hedft_ctx* c, double n4, double* mu4);
/* One statics iteration: STEPRKR when iterrk<=4 else STEPPCR, then derden,
 * fftfw(den), dencg/denalf, poten (BCN4HeDFT.f90:1113-1160). */
int hedft_statics_iteration(hedft_ctx* c, double deltat, int32_t iterrk, double n4,
                            double* mu4, double* mu4err);
/* No counterpart in the reference (its driver reads mu4 back every iteration
 * for the convergence test): niter iterations iterrk = iter0 .. of the same
 * loop as hedft_statics_iteration, queued without a host synchronisation in
 * between; with use_graph the PC iterations before the last replay a captured
 * graph of six (the slot rotation's period).  *mu4, *mu4err, *errpc: those of
 * the last iteration (mu4err against the iteration before it; errpc 0 when the
 * last iteration was an RK one); *elapsed_ms: device time of the iterations
 * (HIP events; any pointer may be NULL).  A caller that tests convergence
 * every k iterations calls this with niter = k. */
int hedft_statics_run(hedft_ctx* c, double deltat, int32_t iter0, int32_t niter, double n4, int32_t use_graph,
                      double* mu4, double* mu4err, double* errpc, double* elapsed_ms);
/* evolo(deltat,mu4,mu4err,n4), He_statics/evolo.f90:5-96: steepest descent
 * psi <- |psi - g + vdt1 (psi - psi1) + vdt2 (psi1 - psi2)| renormalised, with
 * g = deltat (H - mu) psi smoothed iron times by ironing.f90:4-33 (the module
 * values vdt = (0.6, 0.1), iron = 1 are the reference defaults; bulk mode not
 * covered).  psi1 = psi2 = psi are set by hedft_statics_init
 * (BCN4HeDFT.f90:861-862).  *mu4 in: previous mu, out: <psi|H|psi>/n4. */
int hedft_evolo(hedft_ctx* c, double deltat, double vdt1, double vdt2, int32_t iron, double n4,
                double* mu4, double* mu4err);
/* evolox(deltatx,epsx,epsxerr), He_statics/evolox.f90:5-87: one steepest-descent
 * step of the impurity wave function psix of a statics context with a quantum
 * impurity (hedft_set_quantum_impurity on a statics context: real psix, the
 * kernel vq, h2o2mx = h2o2m4 4.002602/m_imp_u as BCN4HeDFT.f90:537) in the
 * potential upotx left by poten: psix <- max(psimin, psix - deltatx g
 * + vdt1 (psix - psi1x) + vdt2 (psi1x - psi2x)) normalised to 1, with g =
 * (H - epsx) psix smoothed ironx times.  *epsx in: previous eigenvalue, out:
 * <psix|H|psix>.  The driver follows it with hedft_statics_refresh. */
int hedft_evolox(hedft_ctx* c, double deltatx, double vdt1, double vdt2, int32_t ironx, double* epsx,
                 double* epsxerr);
/* The driver's part after the steps (BCN4HeDFT.f90:1125-1160): derden,
 * fden = F(den), fdenx = F(psix^2) with an impurity wave function, dencg,
 * denalf, poten.  A statics driver with an impurity runs the helium step
 * (hedft_evolo / hedft_steprkr / hedft_steppcr / hedft_steprkrl /
 * hedft_steppcrl), then hedft_evolox, then this. */
int hedft_statics_refresh(hedft_ctx* c);
/* Statics iteration with the steepest-descent method (Lrkpc = .false.):
 * evolo, then derden, fftfw(den), dencg/denalf, poten (BCN4HeDFT.f90:1114-1160). */
int hedft_statics_iteration_sd(hedft_ctx* c, double deltat, double vdt1, double vdt2, int32_t iron,
                               double n4, double* mu4, double* mu4err);
/* TDDFT start (DFT4He3d.f90:525-601): psi=sqrt(den) (real), den=|psi|^2,
 * timec, potenimp, poten, forceimp, aimp. */
int hedft_tddft_init(hedft_ctx* c);
/* One TDDFT iteration: STEPRK for the first nrk iterations (the driver uses 3,
 * DFT4He3d.f90:677) else STEPPC, then potenimp, poten, forceimp, aimp=F/m.
 * iter counts from 1.  Errors may be NULL (then nothing syncs). */
int hedft_tddft_iteration(hedft_ctx* c, double deltat, int32_t iter, int32_t nrk,
                          double* errHe, double* errimp, double* errvimp);
/* nsteps PC iterations (after the RK start) with the per-step work captured
 * once in a HIP graph and replayed; returns the elapsed device time of the
 * replays measured with HIP events on the context's stream (ms).  The graph
 * holds one period of the slot rotation (6 iterations); a remainder of
 * nsteps % 6 iterations replays a second graph captured from the same slot
 * phase, so no iteration of a graph run is launched one kernel at a time. */
int hedft_tddft_run_pc(hedft_ctx* c, double deltat, int32_t nsteps, int32_t use_graph,
                       double* elapsed_ms);
/* No counterpart in the reference (its time loop has no graphs): captures, on
 * the host only, the graphs that hedft_tddft_run_pc(c, deltat, nsteps, 1)
 * will replay once skip more PC iterations have run (by any entry point), so
 * that the capture does not happen between those iterations and the run.
 * Nothing is launched and the context's state is unchanged. */
int hedft_tddft_prepare_pc(hedft_ctx* c, double deltat, int32_t skip, int32_t nsteps);

/* Batched ensemble (no counterpart in the reference; the reference runs one
 * droplet per process): advances nctx independent TDDFT contexts (any grid,
 * precision and impurity set-up, RK start done, all on one device) by nsteps
 * PC iterations each.  Each context's captured 6-iteration graph is replayed
 * on its own streams, chunk iterations at a time round-robin (chunk rounded up
 * to a multiple of 6; 0 = 6), so the droplets' kernels overlap on the GPU.
 * deltat[i] is the step of context i; *elapsed_ms is the device time of the
 * whole batch (HIP events).  Results equal running each context alone. */
int hedft_batch_run_pc(hedft_ctx* const* ctxs, int32_t nctx, const double* deltat, int32_t nsteps,
                       int32_t chunk, double* elapsed_ms);

/* ---- density / wave-function files (host only; no GPU needed) ----------
 * Text files of the drivers, read as their Fortran list-directed READs do:
 * leading comment lines starting with cchar skipped (titols.f90:12-42), a
 * header record, then the values (x fastest).  Statics density:
 * He_statics/readen.f90:74-84 (written by printout.f90:52-86); TDDFT wave
 * function: Dynamics_TDDFT/readenc.f90:94-110 (written by printoutc.f90:64-74). */
typedef struct hedft_denfile {
  double xmax, ymax, zmax, hx, hy, hz;
  int32_t nx, ny, nz;
  int32_t limp;              /* statics header only (logical) */
  double ximp, yimp, zimp;   /* statics header only */
} hedft_denfile;
/* den may be NULL (header only); otherwise n must be nx*ny*nz. */
int hedft_read_density(const char* path, char cchar, hedft_denfile* hdr, double* den, size_t n);
/* comment: lines written after '#' (may be NULL); values at full precision. */
int hedft_write_density(const char* path, const hedft_denfile* hdr, const double* den, size_t n,
                        const char* comment);
/* psi: n interleaved complex values (2n doubles), may be NULL; rimp/vimp:
 * (nimp,3) with k fastest, up to max_imp impurities. */
int hedft_read_wavefunction(const char* path, char cchar, hedft_denfile* hdr, int32_t* nimp,
                            double* rimp, double* vimp, int32_t max_imp, double* psi, size_t n);
int hedft_write_wavefunction(const char* path, const hedft_denfile* hdr, int32_t nimp,
                             const double* rimp, const double* vimp, const double* psi, size_t n,
                             const char* comment);

/* ---- serving: continuous batching of independent droplet jobs ----------
 * No counterpart in the reference (one droplet per process there).  A job is a
 * Woods-Saxon droplet (rho0 = 0.0218, a = 0.7 A, R = 2.22 N^(1/3), seeded
 * 1e-3 noise, normalised to N) on a cubic grid with one classical impurity
 * (LJ pair potential capped inside 3 A), started as the TDDFT driver does
 * (DFT4He3d.f90:525-601, nrk RK-Gill iterations) and advanced nsteps PC
 * iterations.  With quantum = 1 the impurity is a wave packet instead
 * (Dynamics_quantum_impurities): psix = exp(-|r-r0|^2/(2 width^2) + i k.r),
 * normalised to 1, k = v0/(2 h2o2mimp) (group velocity v0), mass mass_amu,
 * coupled through uimp_k = the 3-D Fourier transform of the same capped LJ
 * potential (potenimpini.f90:19-29; closed-form core + Filon quadrature). */
typedef struct hedft_job {
  double arrival_s;          /* arrival time, seconds after hedft_serve starts */
  int32_t n;                 /* cubic grid edge */
  int32_t precision;         /* HEDFT_PREC_FP64 | HEDFT_PREC_MIXED */
  int32_t nsteps;            /* PC iterations after the RK start */
  int32_t nrk;               /* RK-Gill start iterations (>= 3) */
  double h;                  /* grid step (A); xmax = n h / 2 */
  double n4;                 /* helium atoms */
  double dt;                 /* time step (K^-1) */
  double r0[3];              /* impurity start (A) */
  double v0[3];              /* impurity velocity (A/ps) */
  double mass_amu, eps, sigma; /* impurity mass, LJ epsilon (K) and sigma (A) */
  uint64_t seed;             /* density noise seed */
  int32_t quantum;           /* 0: classical impurity, 1: quantum impurity (wave packet) */
  int32_t pad_;
  double width;              /* quantum: Gaussian width of the packet (A) */
  /* Caller-supplied He-X potential in place of the LJ eps/sigma above (any
   * potential, e.g. one of the reference's V_impur.f90 catalogue evaluated on
   * the host, as its potenimpini does); NULL: LJ.  The arrays are read while
   * the job is admitted and must stay valid until hedft_serve returns. */
  const double* potion;      /* classical: npot values V(i delinter), i < npot, rmax = (npot-1) delinter
                                (potenimpini.f90:31-59, the table hedft_set_impurities takes) */
  int32_t npot;
  int32_t pad2_;
  double delinter;
  const double* uimp_k;      /* quantum: uimp_k on the half spectrum, (n/2+1) n n values
                                (potenimpini.f90:19-29, as hedft_set_quantum_impurity takes) */
} hedft_job;

typedef struct hedft_job_result {
  int32_t status;            /* HEDFT_OK, the job's error, or HEDFT_ESTATE if never run */
  int32_t pad_;
  double admit_s, done_s;    /* seconds since the start of hedft_serve */
  double latency_s;          /* done - arrival */
  double etot, etot4, ekin4, norm; /* energies after the last step */
  double rimp[3];            /* impurity position (quantum: centre of |psix|^2) */
} hedft_job_result;

typedef struct hedft_serve_opts {
  int32_t device;
  int32_t chunk;             /* PC iterations per scheduling quantum (rounded up to 6) */
  double hbm_budget;         /* bytes of device memory the admitted jobs may hold */
  int32_t max_active;        /* 0: limited by the budget only */
  int32_t realtime;          /* 1: admit a job no earlier than its arrival time */
} hedft_serve_opts;

typedef struct hedft_serve_stats {
  double wall_s;             /* host time of the whole loop */
  double steps_per_s;        /* aggregate PC iterations per second */
  double p50_ms_per_1000;    /* latency percentiles, normalised to 1000 steps */
  double p99_ms_per_1000;
  double peak_bytes;         /* largest footprint held at once */
  int32_t max_concurrent;    /* most jobs resident at once */
  int32_t completed;
} hedft_serve_stats;

/* Device bytes a job holds while resident. */
size_t hedft_job_bytes(const hedft_job* job);
/* Runs the request loop until every job has finished; results[i] belongs to
 * jobs[i].  Jobs are admitted in arrival order when their footprint fits the
 * budget; resident jobs advance chunk iterations at a time on their own
 * streams.  stats may be NULL. */
int hedft_serve(const hedft_job* jobs, int32_t njobs, const hedft_serve_opts* opts,
                hedft_job_result* results, hedft_serve_stats* stats);

/* ---- autotuning (no counterpart in the reference) ----------------------
 * Launch-structure choices of one context, picked by timing: x+y FFT launches
 * with at most xy_split_fwd / xy_split_inv arrays use separate row and column
 * passes instead of whole-plane workgroups; no_side runs the independent
 * branches of the mean field on the main stream; prefer_eager replays no
 * graph.  no_side and prefer_eager leave every result bit-identical; the
 * split FFT passes use another radix order, so results agree to rounding. */
typedef struct hedft_tuning {
  int32_t xy_split_fwd, xy_split_inv, no_side, prefer_eager;
  int32_t lap_late;          /* where the next step's x+y Laplacians run: 0 beside
                                the density convolutions, 1 beside the second
                                half of the mean field, 2 the latter only in
                                the mean field that carries the impurity force,
                                3 on the main stream after the density
                                convolutions, 4 on the side stream after
                                grad(den), joined after stage B starts */
  int32_t lap25;             /* TDDFT steppers: > 0 computes the x, y and z
                                second derivatives in one 2.5-D pass over
                                chunks of lap25 planes (the default, 16);
                                0 runs the separate x+y Laplacian pass (kept
                                for comparison).  Fields are identical; errHe
                                sums its blocks in another order. */
  int32_t fused;             /* 1: with lap25 > 0, the mean field's
                                alpha_s divergence (x, y and z) and the pot4
                                assembly are one 2.5-D pass on the main
                                stream, so stage C has no side branch; 0: the
                                x+y divergence on the side stream beside
                                stage C, then a z pass (the default: 2 %
                                faster at 128^3).  Fields are identical. */
  double ms_per_step;        /* measured time of the winning choice (autotune) */
} hedft_tuning;
/* Times trial_steps PC iterations (rounded up to 6, after a 6-iteration
 * warm-up) for each of the 80 combinations of the other fields on this
 * context (lap25 kept) and keeps the fastest.  The context advances by 80
 * (6 + trial) iterations.  hedft_set_tuning returns HEDFT_EGRID for lap25 > 0
 * on a grid the 2.5-D tiles do not divide (nx % 16, ny % 16) or a statics
 * context. */
int hedft_autotune(hedft_ctx* c, double deltat, int32_t trial_steps, hedft_tuning* best);
int hedft_set_tuning(hedft_ctx* c, const hedft_tuning* t);
int hedft_get_tuning(const hedft_ctx* c, hedft_tuning* t);

/* Measurement hook (no counterpart in the reference): nsteps eager PC
 * iterations, exactly as hedft_tddft_run_pc(use_graph=0) runs them, with HIP
 * events on the context's main stream around each launch of the fused
 * H psi + predictor kernel and the fused H psi + corrector kernel; returns
 * their average device durations (ms). */
int hedft_tddft_probe_pc(hedft_ctx* c, double deltat, int32_t nsteps, double* avg_ms_pred,
                         double* avg_ms_corr);

/* Device time (ms, HIP events on the context stream) of the last
 * hedft_tddft_run_pc call. */
int hedft_last_elapsed(hedft_ctx* c, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* HEDFT_H */
