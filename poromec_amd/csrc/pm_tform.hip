// pm_tform.hip -- vertex stresses and Darcy flux of the current state, the
// quantities the pnsf macro prints: tform_pm (Pform.f:433-700) with
// elmt17_pm isw = 3 (Elmlib03.f:2206-2256, darcy_flux Elmlib.f:3178-3200)
// or elmt18_pm isw = 3 (Elmlib03.f:3068-3160) and effective_stress
// (Pform.f:2487-2503).
//
// The reference evaluates every element at its 8 vertices and adds the
// values into the nodes, element by element, then divides by the number of
// elements at each vertex.  A vertex's value from one element needs only
// that element's data at that corner, so here one thread per node gathers
// over the (element, corner) pairs of the node in ascending element order:
// no element work is repeated, no atomics, and every sum runs in the
// reference's order (with -ffp-contract=off the results are the reference's
// bit for bit, given the same u and dp).
#include "pm_device.h"
#include "pm_elemlib.h"
#include "pm_q1.h"

struct TformArgs {
  int n0, n1;          // nodes of this launch
  const int64_t *ptr;  // (nnode + 1) (element, corner) pairs of each node
  const int *src;      // element * 8 + corner, ascending element
  const int *ix;       // (numel, nen) 0-based
  const double *xv;    // (nnodev, 3)
  const int *mat;      // (numel)
  const double *e;     // (numat, PM_PROP)
  const double *u;     // (nnode, 4)
  const double *dp;    // (nnode)
  double gmod, grav[3];
  double *tx, *txe, *txb, *flux;  // (nnode, 6), (nnode, 6), (nnode, 6), (nnode, 3)
  int *err;
  const int *ie;       // element type per material (elmt18_pm: 18)
  const double *vp;    // (numel, 64, 7) integration-point properties
};

// elmt18_pm isw 3 (Elmlib03.f:3068-3160) at vertex v of element el: total
// and Biot stress and Darcy flux at the 8 corner integration points (with
// their properties), extrapolated to the vertex (extrapol_hexa20_v2)
__device__ void corner_elmt18(const TformArgs &A, int el, int v, const int *nodes,
                              const double (*xv)[3], const double *e, double *st, double *sb,
                              double *fl) {
  const double fluid_d = e[6] * 1.0e-06;
  const double fluid_sw = fluid_d * A.gmod;
  double F[15][8];
  for (int j = 0; j < 8; j++) {
    const int q = c_gp_corner[j];
    const double r = c_pg4[q & 3], s = c_pg4[(q >> 2) & 3], t = c_pg4[q >> 4];
    const GpProp G = prop_porosity(A.vp + ((size_t)el * 64 + q) * 7, fluid_sw);
    double xji[3][3];
    const double det = jacobian(xv, r, s, t, xji);
    if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
    for (int k = 0; k < 20; k++) {
      double h, dr, ds, dt, gx, gy, gz;
      dn20(k, r, s, t, h, dr, ds, dt);
      to_global(xji, dr, ds, dt, gx, gy, gz);
      const double *uk = A.u + (size_t)nodes[k] * 4;
      const double u0 = uk[0], u1 = uk[1], u2 = uk[2];
      ux = ux + gx * u0;
      uy = uy + gy * u0;
      uz = uz + gz * u0;
      vx = vx + gx * u1;
      vy = vy + gy * u1;
      vz = vz + gz * u1;
      wx = wx + gx * u2;
      wy = wy + gy * u2;
      wz = wz + gz * u2;
    }
    const double e0 = ux, e1 = vy, e2 = wz, e3 = uy + vx, e4 = vz + wy, e5 = uz + wx;
    double t6[6];
    t6[0] = (e0 + G.b * e1 + G.b * e2) * G.a;
    t6[1] = (G.b * e0 + e1 + G.b * e2) * G.a;
    t6[2] = (G.b * e0 + G.b * e1 + e2) * G.a;
    t6[3] = e3 * G.c * G.a;
    t6[4] = e4 * G.c * G.a;
    t6[5] = e5 * G.c * G.a;
    double dpm = 0.0, pi = 0.0;
    for (int k = 0; k < 8; k++) dpm = dpm + n8(k, r, s, t) * A.dp[nodes[k]];
    for (int k = 0; k < 8; k++) pi = pi + n8(k, r, s, t) * A.u[(size_t)nodes[k] * 4 + 3];
    dpm = G.cbiot * dpm;
    const double pm = G.cbiot * pi;
    t6[0] = t6[0] - dpm;
    t6[1] = t6[1] - dpm;
    t6[2] = t6[2] - dpm;
    for (int c = 0; c < 6; c++) {
      F[c][j] = t6[c];
      F[6 + c][j] = c < 3 ? t6[c] + pm : t6[c];
    }
    double fx = fluid_d * A.grav[0], fy = fluid_d * A.grav[1], fz = fluid_d * A.grav[2];
    for (int k = 0; k < 8; k++) {
      double dr, ds, dt, gx, gy, gz;
      dn8(k, r, s, t, dr, ds, dt);
      to_global(xji, dr, ds, dt, gx, gy, gz);
      const double pk = A.u[(size_t)nodes[k] * 4 + 3];
      fx = fx - gx * pk;
      fy = fy - gy * pk;
      fz = fz - gz * pk;
    }
    F[12][j] = G.perm * fx;
    F[13][j] = G.perm * fy;
    F[14][j] = G.perm * fz;
  }
  for (int c = 0; c < 6; c++) {
    st[c] = extrapol_vertex(F[c], v);
    sb[c] = extrapol_vertex(F[6 + c], v);
  }
  for (int c = 0; c < 3; c++) fl[c] = extrapol_vertex(F[12 + c], v);
}

// Q1/Q1 element (element type 7) at vertex v: the CPU statement's isw 3
// -- D eps - alpha dp and + alpha p, perm (rho_f g - grad p) -- from the
// trilinear fields
__device__ void corner_q1(const TformArgs &A, int v, const int *nodes, const double (*xv)[3],
                          const double *e, double *st, double *sb, double *fl) {
  const double fluid_d = e[6] * 1.0e-06;
  const double perm = e[2] / (fluid_d * A.gmod);
  const double coef_biot = e[4], ps = e[1];
  const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
  const double a = (e[0] * a1) / (a3 * a2), b = ps / a1, c = 0.5 * (a2 / a1);
  double h[8], hr[8], hs[8], ht[8], ji[3][3], hx[8], hy[8], hz[8], u[24], eps[6];
  q1_shape(c_q1_rn[v], c_q1_sn[v], c_q1_tn[v], h, hr, hs, ht);
  const double det = q1_jacobian(xv, hr, hs, ht, ji);
  if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
  for (int k = 0; k < 8; k++) {
    q1_grad(ji, hr[k], hs[k], ht[k], hx[k], hy[k], hz[k]);
    for (int d = 0; d < 3; d++) u[3 * k + d] = A.u[(size_t)nodes[k] * 4 + d];
  }
  q1_strain(hx, hy, hz, u, eps);
  q1_stress(a, b, c, eps, st);
  const int nv = nodes[v];
  const double dpm = coef_biot * A.dp[nv], pm = coef_biot * A.u[(size_t)nv * 4 + 3];
  st[0] = st[0] - dpm;
  st[1] = st[1] - dpm;
  st[2] = st[2] - dpm;
  for (int k = 0; k < 6; k++) sb[k] = st[k] + (k < 3 ? pm : 0.0);
  fl[0] = fluid_d * A.grav[0];
  fl[1] = fluid_d * A.grav[1];
  fl[2] = fluid_d * A.grav[2];
  for (int k = 0; k < 8; k++) {
    const double pk = A.u[(size_t)nodes[k] * 4 + 3];
    fl[0] = fl[0] - hx[k] * pk;
    fl[1] = fl[1] - hy[k] * pk;
    fl[2] = fl[2] - hz[k] * pk;
  }
  fl[0] = fl[0] * perm;
  fl[1] = fl[1] * perm;
  fl[2] = fl[2] * perm;
}

// ET: 17 (elmt17_pm), 18 (elmt17_pm or elmt18_pm by material), 7 (Q1/Q1)
template <int ET>
__global__ __launch_bounds__(256) void k_tform(TformArgs A) {
  // each XCD takes a contiguous range of nodes, so the elements that
  // neighbouring nodes share are gathered through one L2
  const int n = A.n0 + pm_xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= A.n1) return;
  double t[6] = {0, 0, 0, 0, 0, 0}, tb[6] = {0, 0, 0, 0, 0, 0}, f[3] = {0, 0, 0};
  const int64_t p0 = A.ptr[n], p1 = A.ptr[n + 1];
  for (int64_t q = p0; q < p1; q++) {
    const int code = A.src[q];
    const int el = code >> 3, v = code & 7;
    const int *nodes = A.ix + (size_t)el * (ET == 7 ? 8 : 20);
    double xv[8][3];
    for (int k = 0; k < 8; k++)
      for (int d = 0; d < 3; d++) xv[k][d] = A.xv[(size_t)nodes[k] * 3 + d];
    const double *e = A.e + (size_t)A.mat[el] * PM_PROP;
    if (ET == 7 || (ET == 18 && A.ie[A.mat[el]] == 18)) {
      double st[6], sb[6], fl[3];
      if (ET == 7) corner_q1(A, v, nodes, xv, e, st, sb, fl);
      else corner_elmt18(A, el, v, nodes, xv, e, st, sb, fl);
      for (int j = 0; j < 6; j++) {
        t[j] = t[j] + st[j];
        tb[j] = tb[j] + sb[j];
      }
      f[0] = f[0] + fl[0];
      f[1] = f[1] + fl[1];
      f[2] = f[2] + fl[2];
      continue;
    }
    // material constants of elmt17_pm (Elmlib03.f:1962-2000)
    const double fluid_d = e[6] * 1.0e-06;
    const double fluid_sw = fluid_d * A.gmod;
    const double perm = e[2] / fluid_sw;
    const double coef_biot = e[4];
    const double ps = e[1];
    const double a1 = 1.0 - ps, a2 = 1.0 - 2.0 * ps, a3 = 1.0 + ps;
    const double a = (e[0] * a1) / (a3 * a2), b = ps / a1, c = 0.5 * (a2 / a1);
    const double r = c_rn[v], s = c_sn[v], tt = c_tn[v];
    double xji[3][3];
    const double det = jacobian(xv, r, s, tt, xji);
    if (det <= 1.0e-14) atomicExch(A.err, PM_ERR_JACOBIAN);
    // deform3d over the 20 displacement nodes
    double ux = 0, uy = 0, uz = 0, vx = 0, vy = 0, vz = 0, wx = 0, wy = 0, wz = 0;
    for (int k = 0; k < 20; k++) {
      double h, dr, ds, dt, gx, gy, gz;
      dn20(k, r, s, tt, h, dr, ds, dt);
      to_global(xji, dr, ds, dt, gx, gy, gz);
      const double *uk = A.u + (size_t)nodes[k] * 4;
      const double u0 = uk[0], u1 = uk[1], u2 = uk[2];
      ux = ux + gx * u0;
      uy = uy + gy * u0;
      uz = uz + gz * u0;
      vx = vx + gx * u1;
      vy = vy + gy * u1;
      vz = vz + gz * u1;
      wx = wx + gx * u2;
      wy = wy + gy * u2;
      wz = wz + gz * u2;
    }
    const double e0 = ux, e1 = vy, e2 = wz, e3 = uy + vx, e4 = vz + wy, e5 = uz + wx;
    // stress3d, then total = D eps - alpha dp and Biot = total + alpha p
    double st[6];
    st[0] = (e0 + b * e1 + b * e2) * a;
    st[1] = (b * e0 + e1 + b * e2) * a;
    st[2] = (b * e0 + b * e1 + e2) * a;
    st[3] = e3 * c * a;
    st[4] = e4 * c * a;
    st[5] = e5 * c * a;
    const int nv = nodes[v];
    const double dpm = coef_biot * A.dp[nv];
    const double pm = coef_biot * A.u[(size_t)nv * 4 + 3];
    st[0] = st[0] - dpm;
    st[1] = st[1] - dpm;
    st[2] = st[2] - dpm;
    double sb[6];
    sb[0] = st[0] + pm;
    sb[1] = st[1] + pm;
    sb[2] = st[2] + pm;
    sb[3] = st[3];
    sb[4] = st[4];
    sb[5] = st[5];
    // darcy_flux: perm (rho_f g - grad p) from the 8 vertex pressures
    double fx = fluid_d * A.grav[0], fy = fluid_d * A.grav[1], fz = fluid_d * A.grav[2];
    for (int k = 0; k < 8; k++) {
      double dr, ds, dt, gx, gy, gz;
      dn8(k, r, s, tt, dr, ds, dt);
      to_global(xji, dr, ds, dt, gx, gy, gz);
      const double pk = A.u[(size_t)nodes[k] * 4 + 3];
      fx = fx - gx * pk;
      fy = fy - gy * pk;
      fz = fz - gz * pk;
    }
    fx = perm * fx;
    fy = perm * fy;
    fz = perm * fz;
    for (int j = 0; j < 6; j++) {
      t[j] = t[j] + st[j];
      tb[j] = tb[j] + sb[j];
    }
    f[0] = f[0] + fx;
    f[1] = f[1] + fy;
    f[2] = f[2] + fz;
  }
  const int ic = (int)(p1 - p0);
  if (ic > 0) {  // vertices: the average over their elements
    for (int j = 0; j < 6; j++) {
      t[j] = t[j] / ic;
      tb[j] = tb[j] / ic;
    }
    for (int j = 0; j < 3; j++) f[j] = f[j] / ic;
  }
  const double pn = A.u[(size_t)n * 4 + 3];
  for (int j = 0; j < 6; j++) {
    A.tx[(size_t)n * 6 + j] = t[j];
    A.txb[(size_t)n * 6 + j] = tb[j];
    A.txe[(size_t)n * 6 + j] = j < 3 ? t[j] + pn : t[j];
  }
  for (int j = 0; j < 3; j++) A.flux[(size_t)n * 3 + j] = f[j];
}

// the (element, corner) pairs of every node, ascending element
static int ensure_incidence(pm_ctx *c) {
  if (c->d_tf_ptr) return PM_OK;
  const PmHostMesh &m = c->mesh;
  std::vector<int64_t> ptr((size_t)m.nnode + 1, 0);
  for (int el = 0; el < m.numel; el++)
    for (int v = 0; v < 8; v++) {
      const int node = m.ix20[(size_t)el * 21 + v];
      if (node > 0) ptr[node]++;
    }
  for (int n = 0; n < m.nnode; n++) ptr[n + 1] += ptr[n];
  if (ptr[m.nnode] >= ((int64_t)1 << 31) || (int64_t)m.numel * 8 >= ((int64_t)1 << 31))
    return PM_ERR_UNSUPPORTED;
  std::vector<int> src((size_t)std::max<int64_t>(ptr[m.nnode], 1), 0);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int el = 0; el < m.numel; el++)
    for (int v = 0; v < 8; v++) {
      const int node = m.ix20[(size_t)el * 21 + v];
      if (node > 0) src[fill[node - 1]++] = el * 8 + v;
    }
  PM_CHECK_HIP(hipMalloc((void **)&c->d_tf_ptr, sizeof(int64_t) * ptr.size()));
  PM_CHECK_HIP(hipMemcpy(c->d_tf_ptr, ptr.data(), sizeof(int64_t) * ptr.size(),
                         hipMemcpyHostToDevice));
  PM_CHECK_HIP(hipMalloc((void **)&c->d_tf_src, sizeof(int) * src.size()));
  PM_CHECK_HIP(hipMemcpy(c->d_tf_src, src.data(), sizeof(int) * src.size(),
                         hipMemcpyHostToDevice));
  PM_CHECK_HIP(hipMalloc((void **)&c->d_tf_out, sizeof(double) * (size_t)m.nnode * 21));
  return PM_OK;
}

int pm_tform_dev(pm_ctx *c, double *tx, double *txe, double *txb, double *flux) {
  if (!c->have_mesh || !c->have_materials) return PM_ERR_STATE;
  PM_TRY(ensure_incidence(c));
  const int nn = c->mesh.nnode;
  TformArgs A;
  A.ptr = c->d_tf_ptr;
  A.src = c->d_tf_src;
  A.ix = c->d_ix;
  A.xv = c->d_xv;
  A.mat = c->d_mat;
  A.e = c->d_e;
  A.u = c->d_u;
  A.dp = c->d_dp;
  A.gmod = c->gravity_mod;
  for (int d = 0; d < 3; d++) A.grav[d] = c->gravity[d];
  A.tx = c->d_tf_out;
  A.txe = A.tx + (size_t)nn * 6;
  A.txb = A.txe + (size_t)nn * 6;
  A.flux = A.txb + (size_t)nn * 6;
  A.err = c->d_elem_err;
  A.ie = c->d_ie;
  A.vp = c->d_vp;
  const bool v18 = c->vprop[0] != 0;
  PM_CHECK_HIP(hipMemsetAsync(c->d_elem_err, 0, sizeof(int), c->stream));
  // the vertices (nodes 1..nnodev; promotion numbers the mid-edge nodes
  // after them) carry all the element work: their launch is spread over the
  // 8 XCDs on its own (a grid of a multiple of 8 workgroups); the mid-edge
  // nodes only receive zeros and the pressure shift of txe
  const int nv = c->mesh.nnodev;
  A.n0 = 0;
  A.n1 = nv;
  const int gv = (((nv + 255) / 256) + 7) & ~7;
  if (c->mesh.nen == 8) hipLaunchKernelGGL(k_tform<7>, dim3(gv), dim3(256), 0, c->stream, A);
  else if (v18) hipLaunchKernelGGL(k_tform<18>, dim3(gv), dim3(256), 0, c->stream, A);
  else hipLaunchKernelGGL(k_tform<17>, dim3(gv), dim3(256), 0, c->stream, A);
  if (nn > nv) {
    A.n0 = nv;
    A.n1 = nn;
    hipLaunchKernelGGL(k_tform<17>, dim3((nn - nv + 255) / 256), dim3(256), 0, c->stream, A);
  }
  PM_CHECK_HIP(hipGetLastError());
  int err = 0;
  PM_CHECK_HIP(hipMemcpyAsync(&err, c->d_elem_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  const double *outs[4] = {A.tx, A.txe, A.txb, A.flux};
  double *dst[4] = {tx, txe, txb, flux};
  const size_t len[4] = {(size_t)nn * 6, (size_t)nn * 6, (size_t)nn * 6, (size_t)nn * 3};
  for (int i = 0; i < 4; i++)
    if (dst[i])
      PM_CHECK_HIP(hipMemcpyAsync(dst[i], outs[i], sizeof(double) * len[i],
                                  hipMemcpyDeviceToHost, c->stream));
  PM_CHECK_HIP(hipStreamSynchronize(c->stream));
  return err ? err : PM_OK;
}
