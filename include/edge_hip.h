/*
 * edge_hip.h -- C ABI of libedge_hip.so, the MI355X (gfx950) EDGE flow-map
 * time step.
 *
 * Conventions (all entry points):
 *  - Return 0 on success, else a negative status (-1 bad argument, -2 HIP
 *    error, -3 out of memory, -4 no convergence); edge_last_error() gives the
 *    thread-local message.  No exception crosses the ABI.
 *  - Every pointer is a device pointer borrowed for the call, except where a
 *    comment says "host".  `stream` is a hipStream_t (torch's
 *    current_stream().cuda_stream); work is enqueued, not synchronised.
 *  - Grids: nx, ny, nz cells; a field of n components is n contiguous planes
 *    of nx*ny*nz floats, index (i*ny + j)*nz + k, except the smoke: its nc
 *    channels (RGBA, nc = 4) sit together per cell, channel c of cell q at
 *    q*nc + c (one 16-byte load per interpolation tap; 16-byte aligned).  MAC faces: u_x
 *    (nx+1, ny, nz), u_y (nx, ny+1, nz), u_z (nx, ny, nz+1).  A map row T_c
 *    is grad psi_c (3 planes); dT_c holds [d2/dydz, d2/dxdz, d2/dxdy] psi_c.
 *  - All arithmetic is FP32 unless a function says otherwise.
 *
 * The reference is Python + Taichi; each comment names the @ti.kernel (or
 * Python function) an entry point replaces, as file:line in
 * pfm-gatech/EDGE-Flow-Maps.
 */
#ifndef EDGE_HIP_H
#define EDGE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library ------------------------------------------------------------ */
const char* edge_last_error(void);
int edge_abi_version(void);
/* ti.sync() (util/timer.py:19,27) */
int edge_stream_sync(void* stream);

/* ---- Hermite tables and finite differences ------------------------------ */
/* Packs the per-cell Hermite coefficients of a map (3*nx*ny*nz*8 floats,
 * 16-byte aligned; per component a plane of {psi_c, T_c} and one of
 * {dT_c, d2T_c}, 16 bytes per cell each).  kind 0: garm, d2T derived from dT
 * (sim/hermite.py:298-456 + 1258-1287); kind 1: full, d2T given
 * (sim/hermite.py:459-607).  d2T* may be NULL for kind 0. */
int edge_pack_hermite_table(int kind, const float* psi, const float* Tx, const float* Ty, const float* Tz,
                            const float* dTx, const float* dTy, const float* dTz, const float* d2Tx,
                            const float* d2Ty, const float* d2Tz, float* tab, int nx, int ny, int nz, float dx,
                            void* stream);
/* want_dT=1: calc_fd_3_from_grad (sim/hermite.py:1180-1255);
 * want_dT=0: calc_fd_3_from_grad_grad (sim/hermite.py:1258-1287), T* ignored. */
int edge_calc_fd3(int want_dT, const float* Tx, const float* Ty, const float* Tz, float* dTx, float* dTy,
                  float* dTz, float* d2Tx, float* d2Ty, float* d2Tz, int nx, int ny, int nz, float dx,
                  void* stream);
/* The high planes {dT_c, d2T_c} of a garm table, d2T by calc_fd_3_from_grad_grad
 * (sim/hermite.py:1258-1287), whose low planes edge_march_edge_table wrote;
 * after it the table equals edge_pack_hermite_table kind 0 of the same map. */
int edge_pack_table_hi(const float* dTx, const float* dTy, const float* dTz, float* tab, int nx, int ny, int nz,
                       float dx, void* stream);

/* ---- backward-map marches ----------------------------------------------- */
/* advect_psi_garm_ed4_fusion_unroll (sim/hermite.py:1426-1521): reads the old
 * map through its garm table, writes the new psi, T, dT (may alias the fields
 * the table was built from). */
int edge_march_edge(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                    float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy, float* dTz, int nx, int ny,
                    int nz, float dx, float dt, float eps, void* stream);
/* edge_march_edge that also writes the low planes {psi_c, T_c} of the new
 * map's garm table into tab_new (which must not be tab_old); edge_pack_table_hi
 * adds the high planes: the table the map's next use reads (sim/hermite.py:
 * 298-456) without a pass over the stored psi and T. */
int edge_march_edge_table(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                          float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy, float* dTz,
                          float* tab_new, int nx, int ny, int nz, float dx, float dt, float eps, void* stream);
/* advect_ps_hermite_ed8 (sim/hermite.py:1290-1338), fused; tab_old is a full table. */
int edge_march_ed8(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                   float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy, float* dTz, float* d2Tx,
                   float* d2Ty, float* d2Tz, int nx, int ny, int nz, float dx, float dt, float eps, void* stream);
/* advect_ps_garm (sim/hermite.py:1398-1421) up to calc_fd_3_from_grad; full table. */
int edge_march_ge(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                  float* psi, float* Tx, float* Ty, float* Tz, int nx, int ny, int nz, float dx, float dt,
                  void* stream);
/* advect_ps_hermite_ed4_fd_unroll (sim/hermite.py:1375-1395), fused; table
 * built from T by calc_fd_3_from_grad (kind 1 after edge_calc_fd3). */
int edge_march_ed4(const float* ux, const float* uy, const float* uz, const float* tab_old, const float* X,
                   float* psi, float* Tx, float* Ty, float* Tz, int nx, int ny, int nz, float dx, float dt,
                   float eps, void* stream);
/* march_phi_grid -> advance_map on the three face grids (run_edge.py:330-357). */
int edge_advance_map_faces(const float* ux, const float* uy, const float* uz, float* phix, float* phiy, float* phiz,
                           float* Fx, float* Fy, float* Fz, int nx, int ny, int nz, float dx, float dt, void* stream);
/* backtrack_psi_grid (run_efm.py:170-185) in one pass: the face maps (phi, F) from the
 * identity, marched by advance_map through velocity s = 0 .. n-1 with dt dts[s].
 * ux, uy, uz: host arrays of n device pointers (MAC velocity fields); dts: host array. */
int edge_backtrack_faces(const float* const* ux, const float* const* uy, const float* const* uz, const float* dts,
                         int n, float* phix, float* phiy, float* phiz, float* Fx, float* Fy, float* Fz, int nx, int ny,
                         int nz, float dx, void* stream);
/* mode 0: advance_map_0 (sim/hermite.py:786-803); mode 1:
 * advance_map_1_transpose (sim/hermite.py:863-918).  pts: 3 planes of n. */
int edge_advance_points(int mode, const float* ux, const float* uy, const float* uz, float* pts, float* Tx,
                        float* Ty, float* Tz, long n, int nx, int ny, int nz, float dx, float dt, void* stream);
/* No counterpart in the reference (its particles are the inertial ones of
 * sim/lp.py): passive dye particles for the 512x256x256 torus workload.  Each
 * live particle q < num (pos: 3 planes of n) takes the RK4 step of a forward
 * face-map point (run_edge.py:336-357, +dt, B-spline MAC velocity at every
 * stage) and is clamped into [0, n_a dx] per axis (sim/lp.py:158-181). */
int edge_dye_advance(const float* ux, const float* uy, const float* uz, float* pos, long n, long num, int nx, int ny,
                     int nz, float dx, float dt, void* stream);
/* interp_psi_hermite(_garm)(_fd_unroll) (sim/hermite.py:921-1046): compose n
 * points (and their rows, if T* given) with a map table. */
int edge_compose_points(const float* tab, float* pts, float* Tx, float* Ty, float* Tz, long n, int nx, int ny,
                        int nz, float dx, void* stream);
/* calc_ed4_garm (sim/hermite.py:1078-1095) on stored point rows. */
int edge_calc_ed4_garm(const float* T0x, const float* T0y, const float* T0z, const float* T1x, const float* T1y,
                       const float* T1z, const float* T2x, const float* T2y, const float* T2z, const float* T3x,
                       const float* T3y, const float* T3z, float* dTx, float* dTy, float* dTz, long n, float eps,
                       void* stream);
/* np=8: calc_ed8_3 (sim/hermite.py:1098-1151); np=4: calc_ed4_1
 * (sim/hermite.py:1154-1177).  pts: host array of np device pointers. */
int edge_calc_ed(int np, const float* const* pts, float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy,
                 float* dTz, float* d2Tx, float* d2Ty, float* d2Tz, long n, float eps, void* stream);
/* dst = src + (ox, oy, oz): the eps-point resets (sim/hermite.py:704-783). */
int edge_offset_points(const float* src, float* dst, long n, float ox, float oy, float oz, void* stream);
/* Test hooks for the @ti.func interpolators: what 0 interp_(grad_)2 of one
 * field with BL (host float[3]) (util/taichi_utils.py:254-319); 1
 * interp_u_MAC(_grad) (322-341); 2 a map table's value (+rows).  pts: n x 3
 * interleaved. */
int edge_eval_points(int what, const float* f, int d0, int d1, int d2, const float* BL, const float* ux,
                     const float* uy, const float* uz, const float* tab, int nx, int ny, int nz, const float* pts,
                     long n, float dx, float* val, float* grad, void* stream);
/* reset_to_identity_1/2/3 (sim/hermite.py:642-701); NULL outputs are skipped. */
int edge_reset_identity(const float* X, float* psi, float* Tx, float* Ty, float* Tz, float* dTx, float* dTy,
                        float* dTz, float* d2Tx, float* d2Ty, float* d2Tz, int nx, int ny, int nz, void* stream);
/* reset_to_identity on the faces (run_edge.py:306-320). */
int edge_reset_faces(float* phix, float* phiy, float* phiz, float* Fx, float* Fy, float* Fz, int nx, int ny, int nz,
                     float dx, void* stream);
/* center/x/y/z_coords_func (util/taichi_utils.py:23-44): mode 0 cells, 1..3 faces. */
int edge_coords(int mode, float* out, int nx, int ny, int nz, float dx, void* stream);

/* ---- step-loop driver kernels ------------------------------------------- */
/* map_u (run_edge.py:360-373). */
int edge_map_u(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
               const float* Fx, const float* Fy, const float* Fz, const float* phix, const float* phiy,
               const float* phiz, int nx, int ny, int nz, float dx, void* stream);
/* No single counterpart: map_u (run_edge.py:360-373) then add_fields(u1, i, u1, m) and
   scale_field(u1, s, u1) (run_edge.py:671-681): u1 = s (map_u(u0) + m i), the same values
   as the three calls in one pass.  i must not be u1. */
int edge_map_u_add_scale(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                         const float* Fx, const float* Fy, const float* Fz, const float* phix, const float* phiy,
                         const float* phiz, const float* ix, const float* iy, const float* iz, float m, float s,
                         int nx, int ny, int nz, float dx, void* stream);
/* No single counterpart: map_u (run_edge.py:360-373, or map_u_psi 376-417 through the
   map's face values) of the BFECC error u0 into tmp_u, then add_fields(c, tmp_u, err, -1)
   and clamp_u(err, c, tmp_u) (run_edge.py:683-693 with BFECC_clamp):
   u1 = clamp(c - map_u(u0)) to the min and max of c's six neighbours, in one pass (the
   same values; tmp_u and err are not stored).  u1 must differ from u0 and c. */
int edge_map_u_sub_clamp(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                         const float* Fx, const float* Fy, const float* Fz, const float* phix, const float* phiy,
                         const float* phiz, const float* cx, const float* cy, const float* cz, int nx, int ny, int nz,
                         float dx, void* stream);
/* map_u_psi (run_edge.py:376-417); tab: garm table of the cell map. */
int edge_map_u_psi(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                   const float* tab, int nx, int ny, int nz, float dx, void* stream);
/* split_psi_hermite (run_ed8.py:190-223, run_ge.py:184-217); tab: full table. */
int edge_split_psi_hermite(const float* tab, float* phix, float* phiy, float* phiz, float* Fx, float* Fy, float* Fz,
                           int nx, int ny, int nz, float dx, void* stream);
/* mode 0 map_smoke (run_edge.py:420-427), 1 clamp_smoke (469-477); nc 1 or 4. */
int edge_map_smoke(int mode, const float* s0, float* s1, int nc, const float* phix, const float* phiy,
                   const float* phiz, int nx, int ny, int nz, float dx, void* stream);
/* No single counterpart: edge_map_smoke with a third smoke field s2 folded in, the same
   values as the separate calls: mode 0 is map_smoke then add_fields(s1, s2, s1, m) and
   scale_field(s1, s, s1) (run_edge.py:672-682); mode 1 is add_fields(s1, s2, s1, m) then
   clamp_smoke (run_edge.py:688-689); mode 2 is map_smoke(s2, tmp) then mode 1 with tmp
   (run_edge.py:684-689: the error pulled back and the corrected smoke clamped in one
   pass; tmp is not stored).  s0, s1, s2 distinct. */
int edge_map_smoke_add_scale(int mode, const float* s0, float* s1, const float* s2, float m, float s, int nc,
                             const float* phix, const float* phiy, const float* phiz, int nx, int ny, int nz, float dx,
                             void* stream);
/* mode 0 map_smoke_psi (run_edge.py:443-466), 1 clamp_smoke_psi (480-505). */
int edge_map_smoke_psi(int mode, const float* s0, float* s1, int nc, const float* tab, int nx, int ny, int nz,
                       float dx, void* stream);
/* clamp_u (run_edge.py:508-519) on the three face grids. */
int edge_clamp_u(const float* ux, const float* uy, const float* uz, const float* u1x, const float* u1y,
                 const float* u1z, float* u2x, float* u2y, float* u2z, int nx, int ny, int nz, void* stream);
/* No single counterpart: add_fields(u, t, err, -1) then clamp_u(err, u, t) (run_edge.py:685-693,
   508-519) in one pass, in place on t: t = clamp(u - t, min/max of u's six face neighbours).
   Same values as the two calls; err is not written. */
int edge_sub_clamp_u(const float* ux, const float* uy, const float* uz, float* tx, float* ty, float* tz, int nx,
                     int ny, int nz, void* stream);
/* advect_u_semi_lagrangian (sim/hermite.py:1578-1621); dt is the full argument. */
int edge_advect_u_sl(const float* u0x, const float* u0y, const float* u0z, float* u1x, float* u1y, float* u1z,
                     int nx, int ny, int nz, float dx, float dt, void* stream);
/* calc_max_speed (run_edge.py:295-303) into one device float. */
int edge_max_speed(const float* ux, const float* uy, const float* uz, float* out_dev, int nx, int ny, int nz,
                   void* stream);
/* add_fields (util/taichi_utils.py:12-15): dest = f1 + m f2. */
int edge_axpy(const float* f1, const float* f2, float* dest, float m, long n, void* stream);
/* No counterpart in the reference: add_fields then scale_field (run_edge.py:675-681,
 * err = (err - init) * 0.5) in one pass, dest = s * (f1 + m * f2), the same values. */
int edge_axpy_scale(const float* f1, const float* f2, float* dest, float m, float s, long n, void* stream);
/* scale_field (util/taichi_utils.py:6-9). */
int edge_scale(const float* a, float alpha, float* r, long n, void* stream);
/* Taichi Field.copy_from / fill (run_edge.py:621-624, 644). */
int edge_copy(const float* src, float* dst, long n, void* stream);
int edge_fill(float* p, float v, long n, void* stream);
/* add_vortex_ring_and_smoke (sim/init_conditions.py:13-36); c, ex, ey, color: host float[3];
 * vf: 3 planes, smokef: 4 channels per cell. */
int edge_vortex_ring(float w, float rad, float delta, const float* c, const float* ex, const float* ey,
                     const float* pf, float* vf, float* smokef, const float* color, int num_samples, int nx, int ny,
                     int nz, void* stream);
/* get_central_vector (util/taichi_utils.py:95-100): vc = 3 planes of nx*ny*nz cells. */
int edge_get_central_vector(const float* ux, const float* uy, const float* uz, float* vc, int nx, int ny, int nz,
                            void* stream);
/* curl (util/taichi_utils.py:70-92): vf, cf = 3 planes of nx*ny*nz cells, must differ. */
int edge_curl(const float* vf, float* cf, int nx, int ny, int nz, float dx, void* stream);
/* No counterpart in the reference (it runs get_central_vector, curl, numpy's norm and
 * vtk's writer, run_edge.py:720-725, util/io_utils.py:53-75): the device part of a frame
 * output.  wn = |curl u| (nx*ny*nz floats) and sm = the smoke's nc channels per point
 * (nx*ny*nz*nc floats), both in VTK point order (x fastest); nc in 0..4. */
int edge_vtk_frame(const float* ux, const float* uy, const float* uz, const float* smoke, int nc, float* wn,
                   float* sm, int nx, int ny, int nz, float dx, void* stream);
/* split_central_vector (util/taichi_utils.py:103-116). */
int edge_split_central_vector(const float* vc, float* ux, float* uy, float* uz, int nx, int ny, int nz,
                              void* stream);

/* ---- pressure projection (MGPCG) ----------------------------------------- */
/* Solver for an nx*ny*nz cell grid, `levels` multigrid levels (dims divisible
 * by 2^(levels-1), nz even on every level).  mode 0 plain: FP32 storage and
 * the reference V-cycle (sim/mgpcg.py:23-57,147-195); mode 1 mixed: BF16
 * level storage, FP32 arithmetic, dense FP64 coarsest solve; mode 2 dense:
 * FP32 level storage with the dense FP64 coarsest solve.  bt: host
 * int[6] boundary_types ([axis][side], 2 = wall) or NULL for all walls.
 * Replaces MGPCG_3.__init__ (sim/mgpcg.py:366-371) and MGPCG_Solid_3.__init__
 * (sim/mgpcg_solid.py:315-324).  The handle owns its hierarchy. */
int edge_mg_create(int nx, int ny, int nz, int levels, int mode, const int* bt, void** handle);
int edge_mg_destroy(void* handle);
/* Solid cells (sim/mgpcg_solid.py): int32 mask (>0 solid) and boundary
 * velocity (3 planes), borrowed until changed; NULL mask clears. */
int edge_mg_set_solid(void* handle, const int* bm, const float* bvel);
/* MGPCG_3.Poisson / Poisson_with_buffer (sim/mgpcg.py:440-486) and
 * MGPCG_Solid_3.Poisson (sim/mgpcg_solid.py:418-427): u projected in place;
 * p_out, f* optional.  stop_mode 0 = the reference's test (plain rTr <
 * max(abs_tol, rel_tol z0'r0), sim/mgpcg.py:228,254; solid rTr < abs_tol,
 * sim/mgpcg_solid.py:281), 1 = ||r|| <= rel_tol ||b|| (FP64 norms).  x0
 * optional warm start (kept only if ||b - A x0|| <= ||b||).  check_every:
 * iterations between host convergence checks (at most 8), or 0 for batches
 * sized from the solver's last two solves.  stats: host double[9]
 * {iters, rTr, ||b||^2, fluid div sum, corrected, converged, z0'r0,
 * ||b - A x0||^2, x0 used}.  Returns -4 (results written) when max_iters ran
 * out. */
int edge_poisson(void* handle, float* ux, float* uy, float* uz, float* p_out, float* fx, float* fy, float* fz,
                 int max_iters, int stop_mode, double abs_tol, double rel_tol, const float* x0, int check_every,
                 double* stats, void* stream);
/* The right-hand side b (nx*ny*nz) edge_poisson would solve for: u gets the
 * boundary conditions in place (as edge_poisson does first), b = the
 * recentred -div u on fluid cells (sim/mgpcg.py:59-75,169-175,211-217). */
int edge_mg_rhs(void* handle, float* ux, float* uy, float* uz, float* b, void* stream);
/* A x = b with the same solver (MGPCG.solve, sim/mgpcg.py:197-283, after init). */
int edge_mg_solve(void* handle, const float* b, float* x, int max_iters, int stop_mode, double abs_tol,
                  double rel_tol, const float* x0, int check_every, double* stats, void* stream);
/* Test hooks: one V-cycle z = M^-1 r (apply_preconditioner, sim/mgpcg.py:177-195)
 * and Ap = A p (get_Ap, sim/mgpcg.py:119-124). */
int edge_mg_vcycle(void* handle, const float* r, float* z, void* stream);
int edge_mg_apply_A(void* handle, const float* p, float* Ap, void* stream);
/* Replay each batch of CG iterations as captured HIP graphs (enable=1)
 * instead of launching their kernels one by one.  No
 * counterpart in the reference (its CG loop is host-driven, sim/mgpcg.py:234). */
int edge_mg_set_graph(void* handle, int enable);
/* Run each coarse level's V-cycle legs (pre-smoothing + restriction,
 * prolongation + post-smoothing) as one LDS halo-tile launch (enable=1, the
 * default) or as one launch per half-sweep (0).  Same values either way.  No
 * counterpart in the reference (sim/mgpcg.py:177-195 launches per half-sweep). */
int edge_mg_set_tiles(void* handle, int enable);
/* 1 if the mixed solver uses the dense coarsest solve (0: red-black bottom). */
int edge_mg_dense_bottom(void* handle);

/* ---- Learned warm start: 3-level 3D U-Net (16/32/64 channels) ----
 * No counterpart in the reference (a new component).  Activations are
 * channels-last BF16 [voxel][C]; convolution weights [Cout][27][Cin] with
 * tap = (di*3 + dj)*3 + dk.  FP32 gradients and master weights. */
/* 3x3x3 'same' convolution y = relu?(conv(concat(x0, upsample2(x1))) + bias);
 * x1 (half resolution, nearest upsampling) may be NULL with C1 = 0. */
int edge_conv3d_bf16_fwd(const void* x0, int C0, const void* x1, int C1, const void* w, const float* bias, void* y,
                         int Cout, int relu, int n0, int n1, int n2, void* stream);
/* Data gradient of that convolution: g0 [N][S0] += channels < S0, g1 [N/8][Cin-S0]
 * += the others summed over each parent's children; wt = flipped, transposed W. */
int edge_conv3d_bf16_bwd_data(const void* g, int Cg, const void* wt, int Cin, int S0, float* g0, float* g1, int n0,
                              int n1, int n2, void* stream);
/* Weight gradient dw [Cout][27][C0+C1] += sum_v g[v] x(v + tap). */
int edge_conv3d_bf16_bwd_weight(const void* g, int Cout, const void* x0, int C0, const void* x1, int C1, float* dw,
                                int n0, int n1, int n2, void* stream);
/* 2x2x2 average pooling and its gradient. */
int edge_unet_pool(const void* x, void* y, int C, int n0, int n1, int n2, void* stream);
int edge_unet_pool_bwd(const float* gy, float* gx, int C, int n0, int n1, int n2, void* stream);
/* Gradient through ReLU into BF16 (+ bias gradient when db). */
int edge_unet_relu_bwd(const float* gy, const void* y, void* g, float* db, int N, int C, int relu, void* stream);
/* Input features [N][8] from b (scaled by its RMS, returned in scale) and the solid mask. */
int edge_unet_input(const float* b, const int* bm, int N, void* x, float* scale, double* work, void* stream);
/* Output head p = scale (h . w + bias). */
int edge_unet_head(const void* h, const float* w, const float* bias, const float* scale, int N, int C, float* p,
                   void* stream);
/* The head's gradients (gh, dw, db accumulated) for a given dL/dp [N]: used with
 * the energy-norm loss, whose dL/dp is A (p - p*) from the solver's operator. */
int edge_unet_head_bwd(const float* dp, const float* scale, const void* h, const float* w, int N, int C, float* gh,
                       float* dw, float* db, void* stream);
/* Mean-free squared error against a target pressure, its gradient and the head's gradients. */
int edge_unet_loss_head_bwd(const float* p, const float* target, const float* scale, const void* h, const float* w,
                            int N, int C, float* dp, float* gh, float* dw, float* db, double* work, void* stream);
/* BF16 forward (and data-gradient) weights of one layer from its FP32 weights. */
int edge_unet_pack_weights(const float* w, int Cout, int Cin, void* wf, void* wt, void* stream);
/* One Adam step over a flat FP32 parameter vector. */
int edge_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2, float eps, float c1,
              float c2, void* stream);
int edge_zero_f32(float* p, long n, void* stream);

/* ---- Laden particles (sim/lp.py) ----
 * Particle vectors (vel, pos, acc, diff) are 3 planes of n floats (the
 * capacity, laden_particles_max_num); particles q < num are live.  The MAC
 * velocity is the usual three face grids of an nx x ny x nz cell grid. */
/* update_particles_acc (sim/lp.py:6-29); acc = 0 for q >= num. */
int edge_lp_acc(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz, const float* vel,
                const float* pos, float* acc, long n, long num, float dx, float dt, float St, float gravity,
                void* stream);
/* update_particles_velocity (sim/lp.py:32-70). */
int edge_lp_vel(const float* acc, float* vel, const float* pos, long n, long num, float dt, float dx, float Lx,
                float Ly, float Lz, void* stream);
/* calculate_velocity_diff (sim/lp.py:73-97). */
int edge_lp_vel_diff(const float* ux, const float* uy, const float* uz, int nx, int ny, int nz, const float* vel,
                     float* diff, const float* pos, long n, long num, float dx, float St, float density_ratio,
                     float radius_ratio, void* stream);
/* laden_particles_back_with_computing_particle (sim/lp.py:100-156): drag
 * faces zeroed, then the P2G scatter of diff (r_p / radius)^3 (FP32 atomics). */
int edge_lp_p2g(float* dragx, float* dragy, float* dragz, const float* diff, const float* pos, const float* ref_radius,
                long n, long num, float dx, float radius, float Lx, float Ly, float Lz, int nx, int ny, int nz,
                void* stream);
/* No counterpart in the reference: edge_lp_p2g for callers that cannot
 * allocate (code under HIP graph capture, tight stepping loops, harnesses that
 * forbid hipMalloc).  Same arguments and results; the sort keys, cell ranges
 * and sort storage come from a file-scope device array (56 MiB per GPU, held
 * until the process ends) instead of the library's grown workspace.  Limits:
 * min(n, num) <= 1,048,576 particles and at most 4,194,304 tile-padded cells
 * (ceil(n/8)*8 per axis, e.g. 256x128x128); larger calls return EDGE_EINVAL
 * with a message.  Calls that overlap on one GPU (different streams) race. */
int edge_lp_p2g_noalloc(float* dragx, float* dragy, float* dragz, const float* diff, const float* pos,
                        const float* ref_radius, long n, long num, float dx, float radius, float Lx, float Ly,
                        float Lz, int nx, int ny, int nz, void* stream);
/* update_particles_position (sim/lp.py:158-181). */
int edge_lp_pos(const float* vel, float* pos, long n, long num, float dt, float Lx, float Ly, float Lz, void* stream);
/* calculate_viscous_force (sim/lp.py:183-218) on the three face grids. */
int edge_lp_viscous(const float* ux, const float* uy, const float* uz, float* visx, float* visy, float* visz, int nx,
                    int ny, int nz, float dx, float Re, void* stream);
/* calc_max_particle_speed (run_edge.py:258-262): one device float. */
int edge_lp_max_speed(const float* vel, long n, float* out_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EDGE_HIP_H */
