/*
 * pm_march.c -- oracle restatement of the element loop, the CSRC assembly and
 * one backward-Euler time step of the mef `solv` macro.
 *
 * TEST INFRASTRUCTURE ONLY (see pm_oracle.h).
 *
 *   orc_pform          pform_pm sequential branch  Pform.f:1-429 (zeroing :163-168,
 *                      gather :307-381) + csr      Assbly.f:218-304
 *   orc_porosity       porosity_form               Pform.f:969-1176
 *   orc_tform          tform_pm, effective_stress  Pform.f:433-700, 2487-2503
 *   orc_initial_porosity initial_porosity          Pform.f:1331-1376
 *   orc_initial_prop   initial_prop                Pform.f:1434-1540
 *   orc_update_prop    update_prop (isw 10)        Pform.f:1176-1330
 * Elements of materials of type 18 go to elmt18_pm with their points'
 * properties (elmlib_pm, Elmlib.f:77-130).
 *   orc_pload          pload_pm                    Pload.f:1-92
 *   orc_delta_predict  delta_predict_pm            Predict.f:117-158
 *   orc_delta_update   delta_update_pm             Predict.f:159-215
 *   orc_update_res     update_res_v2               Pload.f:483-541
 *   orc_residuo        cal_residuo_pm, criterion 2 Residuo.f:39-136 (incl. the
 *                      resp0 = max(resp0, gresu) update of :119)
 *   orc_step           solv macro body             Main.f:876-1123
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "pm_oracle.h"

#define NEN1 (ORC_NEN + 1)
#define NDF ORC_NDF
#define NST ORC_NST

/* csr (Assbly.f:218-304): b(k) -= p(i); ad(k) += s(i,i); every column l of
 * row k is searched linearly in ja(ia(k):ia(k+1)-1) (columns l > k are not
 * stored and are dropped) */
void orc_assemble_csr(const double *s, const double *p, const int *ld, const int64_t *ia,
                         const int *ja, double *al, double *ad, double *b, int neq, int lhs,
                         int rhs, int nst) {
  for (int i = 0; i < nst; i++) {
    int k = ld[i];
    if (k <= 0 || k > neq) continue;
    if (rhs) b[k - 1] = b[k - 1] - p[i];
    if (lhs) {
      ad[k - 1] = ad[k - 1] + s[(size_t)i * nst + i];
      for (int j = 0; j < nst; j++) {
        int l = ld[j];
        if (l <= 0) continue;
        for (int64_t kk = ia[k - 1]; kk <= ia[k] - 1; kk++)
          if (l == ja[kk - 1]) al[kk - 1] = al[kk - 1] + s[(size_t)j * nst + i];
      }
    }
  }
}

static int model_nen(const orc_model *md) { return md->nen == 8 ? 8 : ORC_NEN; }
int orc_nst(const orc_model *md) { return 3 * model_nen(md) + ORC_NENV; }

/* gathers the element arrays exactly as pform_pm does */
void orc_gather(const orc_model *md, int nel, int *ld, double *ul, double *dpl, double *xl,
                   double *pl) {
  const int *ix = md->ix + (size_t)nel * NEN1;
  int kk = 0;
  const int nen = model_nen(md);
  for (int i = 0; i < nen; i++) {
    int no = ix[i];
    for (int j = 0; j < NDF - 1; j++) {
      ld[kk] = md->id[NDF * (no - 1) + j];
      pl[kk] = 0.0;
      ul[kk] = md->u[NDF * (no - 1) + j];
      kk++;
    }
  }
  for (int i = 0; i < ORC_NENV; i++) {
    int no = ix[i];
    ld[kk] = md->id[NDF * (no - 1) + NDF - 1];
    pl[kk] = 0.0;
    ul[kk] = md->u[NDF * (no - 1) + NDF - 1];
    kk++;
    if (dpl) dpl[i] = md->dp[no - 1];
    for (int j = 0; j < 3; j++) xl[3 * i + j] = md->x[3 * (no - 1) + j];
  }
}

static int elem_type(const orc_model *md, int ma) {
  if (md->nen == 8) return 7;
  return md->ie ? md->ie[ma - 1] : 17;
}

/* elmlib_pm: element nel of material ma */
static int elmlib(orc_model *md, int nel, int ma, const int *iqm, const int *iqh,
                  const double *xl, const double *ul, const double *dpl, double *pl, double *sl,
                  int isw, const orc_env *env) {
  const double *el = md->e + (size_t)(ma - 1) * ORC_PROP;
  if (elem_type(md, ma) == 7) return orc_elmt07(el, iqm, iqh, xl, ul, dpl, pl, sl, isw, env);
  if (elem_type(md, ma) == 18)
    return orc_elmt18(el, iqm, iqh, xl, ul, dpl, pl, sl, md->vpropel + (size_t)nel * 448, isw,
                      env);
  return orc_elmt17(el, iqm, iqh, xl, ul, dpl, pl, sl, isw, env);
}

int orc_pform(orc_model *md, int isw, int lhs, int rhs, double *bvec, const orc_env *env) {
  int ld[NST];
  double ul[NST], pl[NST], dpl[ORC_NENV], xl[3 * ORC_NENV];
  double *sl = malloc(sizeof(double) * NST * NST);
  if (lhs) {
    memset(md->al, 0, sizeof(double) * (size_t)md->nad);
    memset(md->ad, 0, sizeof(double) * (size_t)md->neq);
  }
  int rc = 0;
  for (int nel = 0; nel < md->numel && rc == 0; nel++) {
    orc_gather(md, nel, ld, ul, dpl, xl, pl);
    int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
    rc = elmlib(md, nel, ma, md->eload + 7 * (size_t)nel, md->eloadp + 7 * (size_t)nel, xl, ul,
                dpl, pl, sl, isw, env);
    if (rc) break;
    orc_assemble_csr(sl, pl, ld, md->ia, md->ja, md->al, md->ad, bvec, md->neq, lhs, rhs,
                     orc_nst(md));
  }
  free(sl);
  return rc;
}

void orc_initial_porosity(orc_model *md) {
  int *ic = calloc((size_t)md->nnode, sizeof(int));
  for (int i = 0; i < md->nnode; i++) md->porosity[i] = 0.0;
  for (int nel = 0; nel < md->numel; nel++) {
    int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
    double poro = md->e[(size_t)(ma - 1) * ORC_PROP + 7];
    for (int i = 0; i < ORC_NENV; i++) {
      int no = md->ix[(size_t)nel * NEN1 + i];
      if (no <= 0) continue;
      ic[no - 1]++;
      md->porosity[no - 1] = md->porosity[no - 1] + poro;
    }
  }
  /* a vertex of no element keeps 0 (the reference would divide 0 by 0) */
  for (int i = 0; i < md->nnode; i++)
    if (md->fnno[i] == 1 && ic[i] > 0) md->porosity[i] = md->porosity[i] / ic[i];
  free(ic);
}

int orc_tform(orc_model *md, double *tx, double *txe, double *txb, double *flux,
              const orc_env *env) {
  int ld[NST];
  double ul[NST], pl[NST], dpl[ORC_NENV], xl[3 * ORC_NENV];
  double p[8 * 15];
  int iqz[7] = {0, 0, 0, 0, 0, 0, 0};
  int *ic = calloc((size_t)md->nnode, sizeof(int));
  if (!ic) return -4;
  memset(tx, 0, sizeof(double) * 6 * (size_t)md->nnode);
  memset(txb, 0, sizeof(double) * 6 * (size_t)md->nnode);
  memset(flux, 0, sizeof(double) * 3 * (size_t)md->nnode);
  int rc = 0;
  for (int nel = 0; nel < md->numel; nel++) {
    orc_gather(md, nel, ld, ul, dpl, xl, pl);
    for (int i = 0; i < 8 * 15; i++) p[i] = 0.0;
    int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
    rc = elmlib(md, nel, ma, iqz, iqz, xl, ul, dpl, p, NULL, 3, env);
    if (rc) break;
    for (int i = 0; i < ORC_NENV; i++) {
      int no = md->ix[(size_t)nel * NEN1 + i];
      if (no <= 0) continue;
      ic[no - 1]++;
      for (int j = 0; j < 6; j++) {
        tx[6 * (no - 1) + j] = tx[6 * (no - 1) + j] + p[6 * i + j];
        txb[6 * (no - 1) + j] = txb[6 * (no - 1) + j] + p[6 * i + j + 48];
      }
      for (int j = 0; j < 3; j++) flux[3 * (no - 1) + j] = flux[3 * (no - 1) + j] + p[3 * i + j + 96];
    }
  }
  for (int i = 0; i < md->nnode && rc == 0; i++) {
    if (md->fnno[i] == 1) {
      for (int j = 0; j < 6; j++) {
        tx[6 * i + j] = tx[6 * i + j] / ic[i];
        txb[6 * i + j] = txb[6 * i + j] / ic[i];
      }
      for (int j = 0; j < 3; j++) flux[3 * i + j] = flux[3 * i + j] / ic[i];
    }
    /* effective_stress: Terzaghi's effective stress tx + p on the diagonal */
    double pn = md->u[NDF * i + NDF - 1];
    txe[6 * i] = tx[6 * i] + pn;
    txe[6 * i + 1] = tx[6 * i + 1] + pn;
    txe[6 * i + 2] = tx[6 * i + 2] + pn;
    txe[6 * i + 3] = tx[6 * i + 3];
    txe[6 * i + 4] = tx[6 * i + 4];
    txe[6 * i + 5] = tx[6 * i + 5];
  }
  free(ic);
  return rc;
}

void orc_porosity(orc_model *md, double *dporo, int *ic, const orc_env *env) {
  int ld[NST];
  double ul[NST], pl[NST], xl[3 * ORC_NENV];
  int iqz[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < md->nnode; i++) { ic[i] = 0; dporo[i] = 0.0; }
  for (int nel = 0; nel < md->numel; nel++) {
    orc_gather(md, nel, ld, ul, NULL, xl, pl);
    for (int i = 0; i < ORC_NENV; i++) pl[i] = 0.0;
    int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
    elmlib(md, nel, ma, iqz, iqz, xl, ul, NULL, pl, NULL, 7, env);
    for (int i = 0; i < ORC_NENV; i++) {
      int no = md->ix[(size_t)nel * NEN1 + i];
      if (no <= 0) continue;
      ic[no - 1]++;
      dporo[no - 1] = dporo[no - 1] + pl[i];
    }
  }
  /* with variable properties the elements return the porosity itself
   * (elmt18_pm isw 7), otherwise its change */
  for (int i = 0; i < md->nnode; i++)
    if (md->fnno[i] == 1) {
      if (md->vprop[0]) md->porosity[i] = dporo[i] / ic[i];
      else md->porosity[i] = md->porosity[i] + dporo[i] / ic[i];
    }
}

int orc_update_prop(orc_model *md, const orc_env *env) {
  int ld[NST];
  double ul[NST], pl[NST], xl[3 * ORC_NENV], vpl[448];
  int iqz[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int nel = 0; nel < md->numel; nel++) {
    orc_gather(md, nel, ld, ul, NULL, xl, pl);
    int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
    if (elem_type(md, ma) != 18) continue;  /* elmt17_pm has no isw 10 */
    double *vp = md->vpropel + (size_t)nel * 448;
    memcpy(vpl, vp, sizeof(vpl));
    int rc = orc_elmt18(md->e + (size_t)(ma - 1) * ORC_PROP, iqz, iqz, xl, ul, NULL, pl, NULL,
                        vpl, 10, env);
    if (rc) return rc;
    /* the porosity always, the rest as the conseq flags ask */
    for (int j = 0; j < 64; j++) {
      vp[7 * j] = vpl[7 * j];
      if (md->vprop[1]) vp[7 * j + 1] = vpl[7 * j + 1];
      if (md->vprop[2]) vp[7 * j + 2] = vpl[7 * j + 2];
      if (md->vprop[3])
        for (int k = 3; k < 7; k++) vp[7 * j + k] = vpl[7 * j + k];
    }
  }
  return 0;
}

int orc_initial_prop(orc_model *md, double *e, int numat) {
  const double scale = 1.0e-06;
  for (int ma = 0; ma < numat; ma++) {
    double *em = e + (size_t)ma * ORC_PROP;
    double ym = em[0], ps = em[1], ro = em[5], rof = em[6], poro = em[7];
    double lam = ym * ps / ((1.0 + ps) * (1.0 - 2.0 * ps));
    double mu = (0.5 * ym) / (1.0 + ps);
    double k = lam + (2.0 / 3.0) * mu;
    if (md->vprop[2]) em[5] = (ro - poro * rof) / (1.0 - poro);
    if (md->vprop[3] && orc_ks_and_mus(ym, ps, poro, &k, &mu)) return -1;
    em[0] = k;
    em[1] = mu;
  }
  for (int nel = 0; nel < md->numel; nel++) {
    int ma = md->ix[(size_t)nel * NEN1 + ORC_NEN];
    const double *em = e + (size_t)(ma - 1) * ORC_PROP;
    double k = em[0], mu = em[1], perm = em[2], ro = em[5], rof = em[6], poro = em[7];
    double ibiot = 1 / em[3], cbiot = em[4];
    double *vp = md->vpropel + (size_t)nel * 448;
    for (int i = 0; i < 64; i++) {
      double *v = vp + 7 * i;
      v[0] = poro;
      v[1] = perm;
      v[2] = md->vprop[2] ? (1.0 - poro) * ro + poro * rof : ro;
      v[2] = v[2] * scale;
      if (md->vprop[3]) {
        orc_hashin_shtrikman(poro, k, mu, &v[3], &v[4], &v[5], &v[6]);
      } else {
        v[3] = k;
        v[4] = mu;
        v[5] = ibiot;
        v[6] = cbiot;
      }
    }
  }
  return 0;
}

void orc_pload(orc_model *md, const orc_env *env) {
  double c, vc[3];
  for (int i = 0; i < md->nnode; i++)
    for (int j = 0; j < NDF - 1; j++) {
      int nc = md->nload[NDF * i + j];
      if (nc > 0) {
        c = orc_tload(nc, env->t, md->x + 3 * (size_t)(i < md->nnodev ? i : 0),
                      md->u[NDF * i + j], vc, env);
        md->f[NDF * i + j] = c;
      }
    }
  for (int i = 0; i < md->nnode; i++) {
    if (md->fnno[i] != 1) continue;
    int nc = md->nload[NDF * i + NDF - 1];
    if (nc > 0) {
      c = orc_tload(nc, env->t, md->x + 3 * (size_t)i, md->u[NDF * i + NDF - 1], vc, env);
      md->f[NDF * i + NDF - 1] = c;
    }
  }
  for (int i = 0; i < md->nnode; i++)
    for (int j = 0; j < NDF - 1; j++) {
      int k = md->id[NDF * i + j];
      if (k > 0) md->b0[k - 1] = md->f[NDF * i + j];
      else md->u[NDF * i + j] = md->f[NDF * i + j];
    }
  for (int i = 0; i < md->nnode; i++) {
    if (md->fnno[i] != 1) continue;
    int k = md->id[NDF * i + NDF - 1];
    if (k > 0) md->b0[k - 1] = -env->dt * md->f[NDF * i + NDF - 1];
    else md->u[NDF * i + NDF - 1] = md->f[NDF * i + NDF - 1];
  }
}

void orc_delta_predict(orc_model *md) {
  for (int i = 0; i < md->nnode; i++) {
    for (int j = 0; j < NDF - 1; j++) md->u[NDF * i + j] = md->u[NDF * i + j] - md->u0[NDF * i + j];
    if (md->fnno[i] == 1)
      md->u[NDF * i + NDF - 1] = md->u[NDF * i + NDF - 1] - md->u0[NDF * i + NDF - 1];
  }
}

void orc_delta_update(orc_model *md) {
  for (int i = 0; i < md->nnode; i++)
    for (int j = 0; j < NDF - 1; j++) {
      int k = md->id[NDF * i + j];
      if (k > 0) md->u[NDF * i + j] = md->u[NDF * i + j] + md->x0[k - 1];
    }
  for (int i = 0; i < md->nnode; i++) {
    if (md->fnno[i] != 1) continue;
    int k = md->id[NDF * i + NDF - 1];
    if (k > 0) md->u[NDF * i + NDF - 1] = md->u[NDF * i + NDF - 1] + md->x0[k - 1];
  }
}

void orc_update_res(orc_model *md) {
  for (int i = 0; i < md->nnode; i++) {
    for (int j = 0; j < NDF - 1; j++) {
      md->u[NDF * i + j] = md->u0[NDF * i + j] + md->u[NDF * i + j];
      md->u0[NDF * i + j] = md->u[NDF * i + j];
    }
    if (md->fnno[i] == 1) {
      md->dp[i] = md->dp[i] + md->u[NDF * i + NDF - 1];
      md->u[NDF * i + NDF - 1] = md->u0[NDF * i + NDF - 1] + md->u[NDF * i + NDF - 1];
      md->u0[NDF * i + NDF - 1] = md->u[NDF * i + NDF - 1];
    }
  }
}

/* cal_residuo_pm (Residuo.f:39-136), stop criteria 1 and 2.
 * state[0..1] = resid0 values saved between calls (the reference's SAVE
 * variables res0 / resu0, resp0); state[2..3] receive the current norms.
 * Criterion 1: ||b|| over all equations; criterion 2: ||R_u|| and ||R_p||
 * separately, with resp0 = max(resp0, gresu) at nli = 2 as in Residuo.f:119.
 * Returns istop (2 = converged). */
int orc_residuo(const orc_model *md, const double *b, int nli, double tol, double *state,
                int crit) {
  if (crit == 1) {
    double res = sqrt(orc_dot(b, b, md->neq));
    if (nli == 1) state[0] = res;
    else if (nli == 2) state[0] = fmax(state[0], res);
    state[2] = res;
    state[3] = 0.0;
    return (res / state[0] < tol) ? 2 : 0;
  }
  double resu = 0.0, resp = 0.0;
  for (int i = 0; i < md->nnode; i++)
    for (int j = 0; j < NDF - 1; j++) {
      int neqi = md->id[NDF * i + j];
      if (neqi != 0) resu = resu + b[neqi - 1] * b[neqi - 1];
    }
  for (int i = 0; i < md->nnode; i++) {
    int neqi = md->id[NDF * i + NDF - 1];
    if (neqi != 0) resp = resp + b[neqi - 1] * b[neqi - 1];
  }
  double gresu = sqrt(resu), gresp = sqrt(resp);
  if (nli == 1) {
    state[0] = gresu;
    state[1] = gresp;
  } else if (nli == 2) {
    state[0] = fmax(state[0], gresu);
    state[1] = fmax(state[1], gresu);
  }
  state[2] = gresu;
  state[3] = gresp;
  int istop = 0;
  if (gresu / state[0] < tol) istop++;
  if (gresp / state[1] < tol) istop++;
  return istop;
}

/* One `solv` macro step (Main.f:876-1123) with stop criterion 1 or 2 and
 * the default newton_raphson = false.  env->t must hold t_n; it is advanced here.
 * Returns 0, or a negative code on solver / element failure, or -10 when
 * maxnlit is reached (unless cfg->maxnlit_accept, which ends the step there). */
int orc_step(orc_model *md, const orc_cfg *cfg, orc_env *env, orc_step_info *info) {
  memset(info, 0, sizeof(*info));
  env->t = env->t + env->dt;
  orc_pload(md, env);
  int rc = orc_pform(md, 4, 0, 1, md->b0, env);
  if (rc) return rc;
  orc_delta_predict(md);
  int lhs = 1, i = 1;
  double state[4] = {0, 0, 0, 0};
  for (;;) {
    memcpy(md->b, md->b0, sizeof(double) * (size_t)md->neq);
    rc = orc_pform(md, 2, lhs, 1, md->b, env);
    if (rc) return rc;
    int istop = orc_residuo(md, md->b, i, cfg->tol, state, cfg->stop_crit);
    info->nlit = i;
    if (istop == 2) break;
    int used = 0;
    double res = 0.0;
    int it = orc_solv(md->neq, md->ia, md->ja, md->ad, md->al, md->m, md->b, md->x0, cfg, &used,
                      &res);
    if (info->nsolve < 16) {
      info->iters[info->nsolve] = it;
      info->used[info->nsolve] = used;
      info->resid[info->nsolve] = res;
    }
    info->nsolve++;
    if (it < 0) return it;
    orc_delta_update(md);
    if (i >= cfg->maxnlit) {
      if (cfg->maxnlit_accept) break;
      return -10;
    }
    i++;
    lhs = 0;
  }
  if (md->vprop[0]) {  /* Main.f:1070-1079 */
    rc = orc_update_prop(md, env);
    if (rc) return rc;
  }
  double *dporo = malloc(sizeof(double) * (size_t)md->nnode);
  int *ic = malloc(sizeof(int) * (size_t)md->nnode);
  orc_porosity(md, dporo, ic, env);
  free(dporo);
  free(ic);
  orc_update_res(md);
  return 0;
}
