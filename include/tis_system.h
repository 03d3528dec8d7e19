/*
 * tis_system.h -- plain-C description of one DT15 TIS system (one replica of
 * one molecule set), shared by the HIP library's C-ABI (tis_hip.h) and the
 * CPU oracle.  Only plain pointers and sizes; no torch or HIP types.
 *
 * The fields restate the module globals that the reference's force_sumup /
 * energy_sumup / neighbor read for one replica (var_struct.F90, var_setp.F90,
 * var_simu.F90), with three changes of convention:
 *   - indices are 0-based (Fortran arrays are 1-based);
 *   - per-replica temperature-dependent coefficients (coef_dtrna_st(0,:,grep),
 *     inele%coef(grep), inele%cdist(grep), beta_hbond15) are stored in the
 *     system itself, because one system is one replica;
 *   - coordinates are (3,N) column-major FP64, exactly xyz_mp_rep(:,:,irep).
 *
 * Reference layout: src/var_struct.F90:17-288, src/var_setp.F90:190-200,
 * src/var_setp.F90:563-576, src/var_simu.F90.
 */
#ifndef TIS_SYSTEM_H
#define TIS_SYSTEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MPTYPE codes, src/const_index.F90:226-241 */
enum {
  TIS_MP_VOID = 0, TIS_MP_PRO = 1, TIS_MP_RNA_PHOS = 2, TIS_MP_RNA_BASE = 3,
  TIS_MP_RNA_SUGAR = 4, TIS_MP_ION_MG = 5, TIS_MP_ION_CA2 = 6, TIS_MP_ION_NA = 7,
  TIS_MP_ION_K = 8, TIS_MP_ION_CL = 9, TIS_MP_LIG_X1 = 10
};

/* CLASS codes, src/const_index.F90 (class_type) */
enum { TIS_CLASS_VOID = 0, TIS_CLASS_PRO = 1, TIS_CLASS_RNA = 2,
       TIS_CLASS_LIG = 3, TIS_CLASS_ION = 4 };

/* E_TYPE slots used on this path (src/const_index.F90:99-137).  The parameter
 * constructor there assigns 1..25 to the declared (uncommented) fields in
 * order, so STACK_DTRNA = 16, ..., MAX = 25.  0-based here: slot = E_TYPE - 1.
 * Arrays of energies have TIS_E_MAX entries. */
enum {
  TIS_E_TOTAL = 0, TIS_E_VELO = 1, TIS_E_BOND = 2, TIS_E_BANGLE = 3,
  TIS_E_ELE = 7, TIS_E_STACK = 15, TIS_E_HBOND = 16, TIS_E_EXV = 17,
  TIS_E_TSTACK = 18, TIS_E_THBOND = 19, TIS_E_MAX = 25
};

/* Electrostatics function form, inele%i_function_form (src/force_sumup.F90:235-252) */
enum { TIS_ELE_NONE = -1, TIS_ELE_DH = 0, TIS_ELE_COULOMB = 1, TIS_ELE_EWALD = 2 };

typedef struct tis_params {
  /* DT15 excluded volume (indtrna15, para/rna.para:343-351) */
  double exv_adjust;      /* a = 1.5852 */
  double exv_dist;        /* RNA-RNA D_ij = 3.2 */
  double exv_dist_PS;     /* P-S D_ij = 3.2 */
  double exv_inf;         /* energy HIGH_ENERGY threshold = 0.2 */
  int32_t n_sep_P, n_sep_S, n_sep_B;  /* intra-chain separation 3,3,2 */
  double hb_cutoff_dist;  /* 2.0 */
  /* neighbour lists */
  double neigh_margin;    /* insimu%neigh_margin (8.0) */
  /* electrostatics (inele) */
  int32_t ele_form;       /* TIS_ELE_* */
  int32_t dh_cutoff_type; /* i_DH_cutoff_type: energy cut (cutoff_ele*lambda)^2 if 0 else cutoff_ele^2 */
  double cutoff_ele;      /* inele%cutoff_ele */
  double ele_coef;        /* inele%coef(grep)  = 332.06.../eps(T)  [kcal/mol A] */
  double ele_cdist;       /* inele%cdist(grep) = Debye length [A] */
  double ewld_alpha;      /* Ewald splitting parameter [1/A] */
  int32_t ewld_hmax;      /* |h| <= hmax reciprocal vectors */
  /* hbond valence pass */
  double beta_hb;         /* beta_hbond15 = 1/(kB T) */
  /* periodic box */
  int32_t periodic;       /* inperi%i_periodic */
  double box[3];          /* inperi%psize */
} tis_params;

typedef struct tis_system {
  int32_t nmp;                 /* beads (nmp_all == nmp_real here) */
  int32_t nunit;               /* units (chains, ion groups) */
  const int32_t *unit;         /* [nmp] imp2unit, 0-based */
  const int32_t *mptype;       /* [nmp] imp2type (TIS_MP_*) */
  const int32_t *iclass;       /* [nmp] iclass_mp (TIS_CLASS_*) */
  const double  *exv_rad;      /* [nmp] exv_radius_mp */
  const double  *exv_eps;      /* [nmp] exv_epsilon_mp (= sqrt(epsilon)) */
  const double  *charge;       /* [nmp] coef_charge of the bead, 0 if uncharged */
  const double  *rlist2_unit;  /* [nunit*nunit] rneighbordist2_unit (symmetric) */

  int32_t nbd;                 /* bonds (force_bond.F90) */
  const int32_t *bd_mp;        /* [2*nbd] ibd2mp */
  const double  *bd_nat;       /* [nbd] */
  const double  *bd_coef;      /* [nbd] */

  int32_t nba;                 /* angles (force_bangle.F90) */
  const int32_t *ba_mp;        /* [3*nba] iba2mp */
  const double  *ba_nat;       /* [nba] radians */
  const double  *ba_coef;      /* [nba] */

  int32_t nst;                 /* consecutive stacks (force_dtrna_stack.F90) */
  const int32_t *st_mp;        /* [7*nst] idtrna_st2mp: B1 B2 P1 S1 P2 S2 P3 */
  const double  *st_nat;       /* [3*nst] r0, phi1_0, phi2_0 */
  const double  *st_coef;      /* [4*nst] U0(T), k_r, k_phi1, k_phi2 */

  int32_t ntst;                /* tertiary stacks (force_dtrna_stack_nlocal.F90) */
  const int32_t *tst_mp;       /* [6*ntst] B1 B2 S1 S2 P1next P2next */
  const double  *tst_nat;      /* [6*ntst] */
  const double  *tst_coef;     /* [7*ntst] U0, k1..k6 */
  const int32_t *tst2st;       /* [2*ntst] secondary stack displaced (or -1) */
  const int32_t *tst_excl;     /* [2*ntst] flg_tst_exclusive (0/1) */

  int32_t nhb;                 /* hbond candidates (force_dtrna_hbond15.F90) */
  const int32_t *hb_mp;        /* [6*nhb] B_i B_j S_i S_j P_i+1 P_j+1 (i<j) */
  const double  *hb_nat;       /* [6*nhb] */
  const double  *hb_coef;      /* [7*nhb] U0, k_dist, k_ang1, k_ang2, k_dih1..3 */
  const int32_t *hb_site;      /* [6*nhb] idtrna_hb2hbsite(i,side) at 2*i+side, -1 none */
  const int32_t *hb_tertiary;  /* [nhb] flg_hb_tertiary (0/1) */
  int32_t nhbsite;
  const int32_t *hbsite_valence; /* [nhbsite] nvalence_hbsite */

  /* dynamics (setp_mass_fric.F90) */
  const double *mass;          /* [nmp] cmass_mp */
  const double *fric;          /* [nmp] fric_mp (Langevin: 6 pi eta R / m) */
  const double *radius;        /* [nmp] Stokes radius (BD-HI) */
  double tempk;                /* temperature of this replica */

  tis_params p;
} tis_system;

#ifdef __cplusplus
}
#endif
#endif /* TIS_SYSTEM_H */
