/*
 * spoc_abi.h -- C-ABI of the MI355X (gfx950) SPOC right-hand side.
 *
 * The reference's plugin/operator API for this path is the link-time
 * Fortran subroutine
 *
 *   subroutine cal_var(nf,ne,u0,v0,h0,u1,v1,h1,tseg,pseg,ssit,itm,its,norm,
 *                      nbdl,pbdl,nint,pint,bdl,cip,vect,normt,
 *                      mu0,mb0,fc,beta,w0,z0,g0,hb,time)
 *
 * (src/calvar_lin.f:7-10, identical in src/calvarnl_fs.f:9-12,
 * src/calvarnl_ns.f:9-12, src/calvarnl_inv.f:6-9), selected by which object
 * the makefile links (src/mlin:12, src/mfs:11, src/mns:11, src/minv:11) and
 * reading the operator tables from COMMON /const/ /gauss/ /gaussf/ /gauss2/
 * /gauss3/ /matrices/ (src/shwwnl.f:61-71).  The drop-in Fortran cal_var
 * (spoc_amd/fortran/spoc_calvar.F90) forwards to the functions below.
 *
 * Conventions
 *  - Plain pointers and sizes, no library types.  Host pointers unless a
 *    name says "_dev".
 *  - Modal arrays are column-major (ld, ne): mode m of element e at
 *    a[m + ld*e], rows 0..nm-1 used (the reference uses ld = nnmod = 45).
 *  - Integer connectivity is int32 and 1-based exactly as the reference
 *    stores it (tseg second entry 0 on a wall face).  LOGICAL bdl is int32,
 *    nonzero = true.
 *  - Every function returns 0 on success and a negative SPOC_E* code on
 *    failure; spoc_last_error() gives the message.  The reference has no
 *    status returns (it prints and continues or STOPs); the Fortran shim
 *    turns a nonzero status into STOP.
 *  - One HIP stream per context.  A context is not thread-safe.
 */
#ifndef SPOC_ABI_H
#define SPOC_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* cal_var variants: the four link-time implementations. */
enum {
  SPOC_LIN = 0, /* src/calvar_lin.f   */
  SPOC_FS = 1,  /* src/calvarnl_fs.f  */
  SPOC_NS = 2,  /* src/calvarnl_ns.f  */
  SPOC_INV = 3  /* src/calvarnl_inv.f */
};

/* Arithmetic of the right-hand side. */
enum {
  SPOC_FP64 = 0,     /* plain FP64: reproduces the reference's operation order */
  SPOC_FP32_MFMA = 1 /* mixed precision on the matrix cores: the quadrature of
                        the nonlinear, viscous, Coriolis and forcing terms in
                        FP32 MFMA; the linear gravity-wave core (all of
                        SPOC_LIN), the wall projections, the RK4 state, the
                        mass-matrix factor and the integrals in FP64 */
};

enum {
  SPOC_OK = 0,
  SPOC_EINVAL = -1,   /* bad argument / unsupported (p, ng, ngf) */
  SPOC_EHIP = -2,     /* HIP runtime error */
  SPOC_ESTATE = -3,   /* call out of order (e.g. no mesh yet) */
  SPOC_ESINGULAR = -4, /* mass matrix not positive definite (reference only
                          prints "pb avec matrice", src/shwwnl.f:153-156) */
  SPOC_EMESH = -5,      /* mesh rejected: a triangle with two wall faces or a
                          wall face that is not its edge 1 (src/cmal_s.f:222-277),
                          or inconsistent connectivity */
  SPOC_EFILE = -6,      /* file missing, unreadable or not in the expected format */
  SPOC_ENAN = -7,       /* the state is no longer finite (the driver's testnan,
                           src/shwwnl.f:426-445, prints and STOPs); the results
                           are still written and can be downloaded */
  SPOC_ENOMEM = -8      /* host memory, or a host thread, could not be had */
};

typedef struct spoc_ctx spoc_ctx;

const char *spoc_last_error(void);
int spoc_version(void);

/*
 * Operator tables, built from the quadrature rule exactly as the reference
 * driver does (src/shwwnl.f:127-156: cal_p2, cal_pico, cal_mat, cal_matl,
 * DPOTRF).  For callers that do not hold the Fortran COMMONs.  Output
 * layouts are the COMMON layouts with caller-chosen leading dimensions:
 * pi2(ld_g, nm), pico(ld_f, ld_m, 3), matrices (ld_m, ld_m).  amm/amml are
 * returned factored (upper Cholesky factor in the upper triangle, original
 * values below), ammo unfactored.  Returns SPOC_ESINGULAR if DPOTRF fails.
 */
int spoc_tables_build(int nc, int ng, int ngf, const double *wi, const double *xi,
                      const double *yi, const double *wif, const double *tif, int ld_g,
                      int ld_f, int ld_m, double *pi2, double *pico, double *amm, double *amml,
                      double *agx, double *agy, double *ammo);

/*
 * Context: quadrature data and the operator tables in their COMMON layout
 * (replaces /const/ /gauss/ /gaussf/ /gauss2/ /gauss3/ /matrices/,
 * src/shwwnl.f:61-71).  amm and amml must be the factored matrices the
 * driver's DPOTRF left (src/shwwnl.f:153-156).  device = HIP device ordinal.
 */
int spoc_ctx_create(spoc_ctx **ctx, int device, int nc, int ng, int ngf, const double *wi,
                    const double *wif, const double *pi2, int ld_g, const double *pico,
                    int ld_f, int ld_m, const double *amm, const double *amml,
                    const double *agx, const double *agy, const double *ammo);
int spoc_ctx_destroy(spoc_ctx *ctx);

/*
 * Mesh operands of cal_var (computed by the driver: pointer_mesh, cal_its,
 * cal_rotnf, calpre; src/shwwnl.f:162-168): itm/its (3,ne), tseg/ssit (2,nf),
 * norm (2,nf), vect (5,ne), normt (2,ne), bdl (ne).  in (3,ne), xgr, ygr (nn)
 * are used only by spoc_energy (src/energy.f:37-44 recomputes det from the
 * nodes); pass NULL to skip.  Replaces any previous mesh (AMR remeshing).
 */
int spoc_mesh_set(spoc_ctx *ctx, int ne, int nf, const int *itm, const int *its,
                  const int *tseg, const int *ssit, const double *norm, const double *vect,
                  const double *normt, const int *bdl, int nn, const int *in,
                  const double *xgr, const double *ygr);

/*
 * Ensemble: nmem independent meshes run together as one union mesh, so that
 * every kernel launch covers all members.  Arrays are the members' own
 * arrays (as for spoc_mesh_set, 1-based within each member) concatenated
 * member after member; ne_m, nf_m, nn_m give each member's sizes (nn_m,
 * in, xgr, ygr may be NULL).  The library shifts the numbering.  Fields and
 * state are then passed for the union: (ld, sum ne_m) arrays, members
 * concatenated; the linear variant's depth hb(1,1) is taken per member.
 * A member's results equal those of running it alone (the kernels treat
 * each element column independently).  The members share g0, mu0, mb0 and
 * dt.  No reference counterpart (the reference runs members serially).
 */
int spoc_ensemble_mesh_set(spoc_ctx *ctx, int nmem, const int *ne_m, const int *nf_m,
                           const int *nn_m, const int *itm, const int *its, const int *tseg,
                           const int *ssit, const double *norm, const double *vect,
                           const double *normt, const int *bdl, const int *in,
                           const double *xgr, const double *ygr);

/* Per-member integrals of the device-resident state (src/energy.f:3-96):
 * arrays of nmem values, each summed in the member's own element order. */
int spoc_ensemble_energy(spoc_ctx *ctx, int nmem, double *ke, double *pe, double *mass);

/*
 * Element numbering inside the library (no reference counterpart).  The
 * kernels work on blocks of 16 consecutive elements and gather neighbours'
 * traces, so they are fastest when consecutive elements are close in space;
 * the caller's numbering need not be (src/cmal_s.f:164-208 numbers as the
 * mesh file lists triangles, and remeshing reorders).  SPOC_ORDER_HILBERT
 * and SPOC_ORDER_ZCURVE renumber elements along that space-filling curve of
 * their centroids inside each ensemble member (centroids reconstructed from
 * vect and the face connectivity; no node arrays needed) and faces in order
 * of first appearance; SPOC_ORDER_INPUT keeps the caller's numbering;
 * SPOC_ORDER_AUTO (the default) takes the Hilbert numbering unless the
 * caller's already puts at least as many face neighbours in the same block
 * of 16 elements.  Every array crossing the C-ABI stays in the caller's
 * numbering, every result is the same bit for bit, and energies are summed
 * in the caller's element order.  Applies from the next spoc_mesh_set /
 * spoc_ensemble_mesh_set.
 */
enum { SPOC_ORDER_INPUT = 0, SPOC_ORDER_HILBERT = 1, SPOC_ORDER_ZCURVE = 2, SPOC_ORDER_AUTO = 3 };
int spoc_set_ordering(spoc_ctx *ctx, int mode);

/*
 * The internal numbering that mode (SPOC_ORDER_HILBERT or SPOC_ORDER_ZCURVE)
 * gives a mesh (host side, no device work): arrays as for
 * spoc_ensemble_mesh_set (nmem = 1 for one mesh); perm (sum ne_m) receives,
 * for each internal position, the caller's element (1-based, members'
 * numbers shifted as in the union mesh).
 */
int spoc_mesh_order(int mode, int nmem, const int *ne_m, const int *nf_m, const int *itm,
                    const int *tseg, const int *ssit, const double *vect, int *perm);

/*
 * Per-run physics operands of cal_var: fc (Coriolis, modal), w0/z0 (wind
 * stress), hb (resting depth), all (ld, ne); mu0, mb0, g0 as in oc.inp.
 * fc/w0/z0 may be NULL for SPOC_LIN (unused by src/calvar_lin.f).
 */
int spoc_fields_set(spoc_ctx *ctx, int ld, const double *fc, const double *w0, const double *z0,
                    const double *hb, double mu0, double mb0, double g0);

/*
 * One cal_var evaluation on host arrays: (u0,v0,h0) -> tendencies (u1,v1,h1),
 * rows 0..nm-1 of each column written, as the reference does.  This is the
 * call the Fortran drop-in makes; it copies state over PCIe each call.
 */
int spoc_rhs(spoc_ctx *ctx, int variant, int precision, int ld, const double *u0,
             const double *v0, const double *h0, double *u1, double *v1, double *h1);

/* Device-resident state for the fused RK4 fast path. */
int spoc_state_upload(spoc_ctx *ctx, int ld, const double *u0, const double *v0,
                      const double *h0);
int spoc_state_download(spoc_ctx *ctx, int ld, double *u0, double *v0, double *h0);

/*
 * nsteps classic RK4 steps of the driver loop (src/shwwnl.f:233-284, same in
 * src/shwwnl_adap.f:302-353) on the device-resident state.  SPOC_FP64
 * reproduces the reference's arithmetic order, including
 * u0 + (dt/6)*(((k1 + 2k2) + 2k3) + k4).  Returns SPOC_ENAN if the state
 * holds a NaN or an infinity after the last step (the driver's testnan,
 * src/shwwnl.f:297, checks the kinetic energy every per2 steps; this checks
 * every coefficient once per call).  At most 512 kernel dispatches are
 * queued ahead of the host: the call waits for earlier steps as it goes.
 */
int spoc_rk4_steps(spoc_ctx *ctx, int variant, int precision, double dt, int nsteps);

/*
 * Replay the RK4 step of spoc_rk4_steps from a HIP graph (on = 1, the
 * default): the step's launches are captured once per (variant, precision,
 * dt, mesh, fields) and replayed nsteps times, which removes the host launch
 * cost that dominates small meshes.  Used when nsteps >= 2 and profiling is
 * off; results are identical either way.  No reference counterpart.
 */
int spoc_set_graph(spoc_ctx *ctx, int on);

/*
 * Run the linear variant's RK4 steps on the FP32-MFMA path in one
 * persistent launch (on = 1, the default) on small meshes, where that is
 * faster than a launch per stage: at most 1,536 elements at p <= 3, 1,024
 * at p = 4, 768 at p = 5 (none above), with the central flux.  The launch
 * keeps every element block on its wave for all the steps, and before each
 * stage a block waits only for its neighbouring blocks, in place of four
 * launches per step (with profiling on, spoc_last_timing names it
 * "lin_persistent").  Results are identical bit for bit to the
 * launch-per-stage path.  The wait is bounded: if it runs out,
 * spoc_rk4_steps returns SPOC_EHIP.  on = 2 runs the launch with its abort
 * word raised from the start, so that every wave leaves at its first wait
 * and the call returns SPOC_EHIP: the failure path, forced from the host,
 * for tests.  No reference counterpart (the reference's RK4 loop,
 * src/shwwnl.f:233-284, calls the right-hand side once per stage).
 */
int spoc_set_persistent(spoc_ctx *ctx, int on);

/*
 * Discrete adjoint (no reference counterpart).  With R the right-hand side
 * in the reference's modal coordinates (what spoc_rhs computes), the
 * adjoint of one evaluation is (ru,rv,rh) = J_R(u0,v0,h0)^T (au,av,ah)
 * under the plain dot product of the (nm, ne) coefficient arrays, and, for
 * the nonlinear variants, rhb = (dR/dhb)^T (au,av,ah) (the bathymetry hb
 * enters through the total depth eta + hb; NULL to skip; zero for SPOC_LIN,
 * whose depth is the constant hb(1,1)).  The state (u0,v0,h0) is the
 * linearisation point; SPOC_LIN ignores it (may be NULL).
 *
 * spoc_rk4_adjoint_steps runs nsteps steps of the transposed RK4 recursion
 * backwards on the device-resident adjoint state: spoc_adjoint_upload sets
 * the adjoint of the state after nsteps steps, spoc_adjoint_download
 * returns that of the initial state, spoc_adjoint_hb_download the gradient
 * with respect to hb.  For the nonlinear variants the trajectory starts at
 * the device-resident state (spoc_state_upload; left unchanged) and is
 * computed by the mixed-precision forward path.  spoc_set_checkpoint(ctx,
 * every): every >= 1 keeps checkpoints every `every` steps and recomputes
 * the steps between them; every = 0 (the default) keeps every step's four
 * stage inputs from the forward sweep when they fit in half the device's
 * free memory, so that nothing is recomputed, and takes checkpoints every 50
 * steps otherwise.  Results do not depend on the choice.  Everything runs
 * in FP64 on MFMA.
 */
int spoc_rhs_adjoint(spoc_ctx *ctx, int variant, int ld, const double *u0, const double *v0,
                     const double *h0, const double *au, const double *av, const double *ah,
                     double *ru, double *rv, double *rh, double *rhb);
int spoc_adjoint_upload(spoc_ctx *ctx, int ld, const double *au, const double *av,
                        const double *ah);
int spoc_adjoint_download(spoc_ctx *ctx, int ld, double *au, double *av, double *ah);
int spoc_adjoint_hb_download(spoc_ctx *ctx, int ld, double *ahb);
int spoc_set_checkpoint(spoc_ctx *ctx, int every);
int spoc_rk4_adjoint_steps(spoc_ctx *ctx, int variant, double dt, int nsteps);

/*
 * Autotuner of the mixed-precision element kernels (no reference
 * counterpart).  Each element kernel (lin_kernel, or bc_kernel and
 * eg_kernel) is compiled for every polynomial degree in 3 forms (register
 * budget as resident waves per SIMD, and for eg_kernel whether a block of
 * 16 elements is split over two waves) and can run with 2, 4 or 8 waves of
 * 16 elements per workgroup.  A tile is coded as waves + 16 x form (4 is
 * form 0 at 4 waves).  spoc_autotune times all 9 pairs of every kernel on
 * the current mesh and keeps the fastest for `variant` (tiles[0] = lin/bc,
 * tiles[1] = eg, 0 for lin).  It also runs by itself on the first
 * mixed-precision call after a mesh change.  spoc_set_tiles forces a code
 * (0 = automatic) and turns the automatic tuning on or off.  Results do not
 * depend on the tiles.
 */
int spoc_autotune(spoc_ctx *ctx, int variant, int *tiles);
int spoc_set_tiles(spoc_ctx *ctx, int variant, int tile0, int tile1, int autotune);

/*
 * Face flux of the FP32-MFMA path's gravity-wave core.  alpha = 0 (the
 * default) is the reference's central flux (src/calflu.f:41-62).  alpha > 0
 * adds a local Lax-Friedrichs (Rusanov) penalty: the normal flux of eta,
 * u and v gains (c/2)(q_own - q_neighbour) with c = alpha sqrt(g0 hh0), the
 * linear gravity-wave speed (alpha = 1: Rusanov).  It has no counterpart in
 * the reference: it dissipates energy at the jumps, conserves mass, and is
 * zero at walls (walls stay the projection of the reference).  The plain
 * FP64 path and the adjoint support the central flux only (SPOC_EINVAL).
 */
int spoc_set_flux(spoc_ctx *ctx, double alpha);

/* Wait for all work queued on the context's stream. */
int spoc_sync(spoc_ctx *ctx);

/*
 * Mesh operands of cal_var from the nodes and triangles, host side, as the
 * reference driver computes them (src/shwwnl.f:162-174): face numbering
 * pointer_mesh (src/cmal_s.f:164-208), check_mesh (src/cmal_s.f:222-277,
 * skipped when check = 0), cal_rotnf (src/mat1.f:7-23), cal_its
 * (src/mat1.f:28-64), calpre's geometry and wall flags (src/shwwnl.f:350-422)
 * and coriolis (src/mat1.f:69-83).  in (3, ne) 1-based CCW triangles; x, y
 * (nn).  Outputs (caller-allocated; nf <= 3 ne): *nf; itm, its (3, ne);
 * tseg, pseg, ssit (2, nf); norm (2, nf); vect (5, ne); normt (2, ne);
 * bdl (ne, 0/1); fc (ld, ne) with ld >= nm, or NULL; deta (ne) or NULL.
 * No device work.  SPOC_EMESH if the mesh is rejected.
 */
int spoc_mesh_build(int nn, int ne, const int *in, const double *x, const double *y, int nc,
                    double f0, double beta, int ld, int check, int *nf, int *itm, int *its,
                    int *tseg, int *pseg, int *ssit, double *norm, double *vect, double *normt,
                    int *bdl, double *fc, double *deta);

/*
 * The CM tool (src/chmal.f:52-142), in place on in (3, ne): flip the
 * diagonal of triangles with two wall faces, then rotate each wall
 * triangle's local numbering so that its wall face is its edge 1.  Host
 * side.
 */
int spoc_mesh_changemaille(int nn, int ne, int *in);

/*
 * The driver's state files init.bin / temp.bin, host side: 7 Fortran
 * unformatted sequential records (4-byte little-endian byte counts around
 * each): time, then u, v, h, w, z, hb as (nm, ne) element-major
 * (src/shwwnl.f:180-187 reads them, src/shwwnl_adap.f:703-724 writes them).
 * Arrays are (ld, ne) column-major, ld >= nm; on read any of the six may be
 * NULL to skip it.  Records of 2 GiB or more are not supported (SPOC_EFILE).
 */
int spoc_state_file_read(const char *path, int nm, int ne, int ld, double *time, double *u0,
                         double *v0, double *h0, double *w0, double *z0, double *hb);
int spoc_state_file_write(const char *path, int nm, int ne, int ld, double time,
                          const double *u0, const double *v0, const double *h0,
                          const double *w0, const double *z0, const double *hb);

/*
 * The mesh file cm.dat, host side (src/cmal_s.f:7-48 reads it, 56-78 writes
 * it): "nn ne", then nn lines "i x y", then ne lines "e n1 n2 n3",
 * list-directed.  Read multiplies the coordinates by scx, scy as read_mesh
 * does; write divides by them and prints 17 significant digits, which read
 * back to the same doubles.  spoc_mesh_file_size reads the header only.
 */
int spoc_mesh_file_size(const char *path, int *nn, int *ne);
int spoc_mesh_file_read(const char *path, int nn, int ne, double scx, double scy, double *x,
                        double *y, int *in);
int spoc_mesh_file_write(const char *path, int nn, int ne, const double *x, const double *y,
                         const int *in, double scx, double scy);

/*
 * Fields carried onto a new mesh by the adaptive driver, host side, bit for
 * bit as the reference computes them.  Tables as spoc_tables_build returns
 * them: the triangle rule wi, xi, yi (ng), pi2 (ld_g, nm), amm (ld_m, nm)
 * the mass matrix's upper Cholesky factor.
 *
 * spoc_interpol (interpol + interlin, src/adapt.f:1662-1824): nfield modal
 * fields u[f] (ld, ne) of the old mesh (nodes xgr, ygr (nn); in, itm (3, ne);
 * tseg, pseg (2, nf)) onto the new mesh (xgrnew, ygrnew (nnnew); innew
 * (3, nenew)):
 * element i with tinter[i] = t > 0 copies old element t, the others are the
 * L2 projection of the old field sampled at their quadrature points, each
 * located by the reference's walk (from the previous point's element; a
 * point outside the mesh is moved onto the wall edge it left by).  The walk
 * does not depend on the field, so one call carries u, v, h together.
 * SPOC_EMESH if a walk does not end; SPOC_EINVAL, before any walk, if an
 * index array points outside its range.
 *
 * spoc_calh2 (src/adapt.f:303-373): L2 projection onto the mesh (nodes
 * xgr, ygr (nn); in (3, ne)) of the function ff(x, y, user) (the driver's
 * wind and bathymetry), or h0's element tinter[i] where tinter[i] > 0 (h0
 * is (ld, ne0); may be NULL when every tinter[i] is 0).
 */
int spoc_interpol(int nc, int ng, const double *wi, const double *xi, const double *yi,
                  const double *pi2, int ld_g, const double *amm, int ld_m, int nn, int ne,
                  int nf, const double *xgr, const double *ygr, const int *in, const int *itm,
                  const int *tseg, const int *pseg, int nnnew, int nenew, const double *xgrnew,
                  const double *ygrnew, const int *innew, const int *tinter, int nfield, int ld,
                  const double *const *u, double *const *unew);
int spoc_calh2(int nc, int ng, const double *wi, const double *xi, const double *yi,
               const double *pi2, int ld_g, const double *amm, int ld_m,
               double (*ff)(double x, double y, void *user), void *user, int nn, int ne,
               const double *xgr, const double *ygr, const int *in, const int *tinter, int ne0,
               int ld, const double *h0, double *h1);

/*
 * Mesh refinement and coarsening of the adaptive driver, host side, equal
 * to the reference's meshes bit for bit (node coordinates, node and
 * triangle numbering).  A spoc_amr holds the refinement hierarchy, the
 * reference's COMMON /adap_const/, /adap_int/, /adap_vec/, /adap_log/.
 *
 * spoc_amr_create (init_adap, src/adapt.f:380-449): the hierarchy of a
 * mesh of nn nodes (xgr, ygr) and ne triangles (in (3, ne), 1-based), with
 * its faces (itm (3, ne), tseg (2, nf)) as spoc_mesh_build numbers them.
 *
 * spoc_amr_adapt: one adaptation as global_refinment runs it between
 * cal_conv and the interpolation (src/adapt.f:89-119): the flags of the
 * current mesh (fconv: refine, fderaff: coarsen; ne values each, 0/1, as
 * spoc_cal_conv returns them) onto the hierarchy, deraffiner (1393-1583),
 * adapmaille (532-788), protridouble (789-998), elimpptri (999-1121).  The
 * reference runs it when cal_conv's conv is true.  *nn_new, *ne_new: the
 * new mesh's sizes.  The per-element fderaff is read by hierarchy number,
 * as the reference's call does (src/adapt.f:117): numbers past ne count as
 * false.  After an error other than SPOC_EINVAL for the arguments (e.g.
 * SPOC_ENOMEM midway) the hierarchy is not usable: destroy it.
 *
 * spoc_amr_mesh: the current mesh, xgr, ygr (nn_new), in (3, ne_new),
 * 1-based, CCW when the input mesh was; tabsol (ne_new, may be NULL) the
 * hierarchy number of each triangle (past the hierarchy: green halves).
 * spoc_amr_levels: the refinement level of each triangle (a green half at
 * its parent's).
 *
 * spoc_amr_inter_only_new (inter_only_new, src/adapt.f:1233-1258):
 * tinter (nenew) for spoc_interpol: the old triangle that reappears
 * unchanged as each new triangle, or 0.
 */
typedef struct spoc_amr spoc_amr;
int spoc_amr_create(spoc_amr **amr, int nn, int ne, int nf, const double *xgr, const double *ygr,
                    const int *in, const int *tseg, const int *itm);
int spoc_amr_adapt(spoc_amr *amr, int ne, const int *fconv, const int *fderaff, int *nn_new,
                   int *ne_new);
int spoc_amr_mesh(const spoc_amr *amr, double *xgr, double *ygr, int *in, int *tabsol);
int spoc_amr_levels(const spoc_amr *amr, int *level);
void spoc_amr_destroy(spoc_amr *amr);
int spoc_amr_inter_only_new(int ne, const int *in, int nenew, const int *innew, int *tinter);

/*
 * Refinement indicator of the device-resident state (checkconv,
 * src/adapt.f:1260-1332), FP64, bit-identical to the reference: for field
 * 0 (u), 1 (v) or 2 (h), amax[e] (ne values) = the largest |difference| of
 * the two traces at the ngf points of e's faces that have a second
 * element, divided by max |a(1,.)| of the field over e's mesh (ensemble:
 * over e's member), or 0 when that maximum is <= 1e-6.
 */
int spoc_checkconv(spoc_ctx *ctx, int field, double *amax);

/*
 * Refinement decisions (cal_conv, src/adapt.f:1122-1192) on the
 * device-resident state, from the three indicators of spoc_checkconv:
 * fconv[e] = 1 if some indicator of e exceeds fac1, fderaff[e] = 1 if all
 * three are below fac2 (ne values each).  Per mesh (ensemble: per member,
 * nmem values each): deraff = some fderaff, supmax = the largest indicator,
 * timedir = (supmax <= fac3) (0 asks the driver to roll back), conv =
 * (supmax > fac1 or deraff).  LOGICALs as 0/1.  supmax and amax3 (3 x ne,
 * the indicators of u, v, h) may be NULL.
 */
int spoc_cal_conv(spoc_ctx *ctx, double fac1, double fac2, double fac3, int *fconv,
                  int *fderaff, int *conv, int *deraff, int *timedir, double *supmax,
                  double *amax3);

/*
 * Global integrals of the device-resident state (src/energy.f:3-96), FP64:
 * kinetic 1/2 int (h+H)|u|^2, potential 1/2 g int h^2, mass int h.  Needs the
 * node arrays passed to spoc_mesh_set.  The three values are written, and
 * SPOC_ENAN returned, when the kinetic energy is NaN (the test of testnan,
 * src/shwwnl.f:426-445).
 */
int spoc_energy(spoc_ctx *ctx, double *ke, double *pe, double *mass);

/*
 * spoc_energy without allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops).  Queues the integrals of the device-resident state
 * on the context's stream and returns without waiting: out_dev (device
 * memory, 3 x nmem doubles, nmem = 1 for one mesh) receives ke, pe, mass per
 * ensemble member, bit-identical to spoc_energy / spoc_ensemble_energy,
 * once the stream has run (spoc_sync).  The per-element partial integrals
 * live in a fixed device array of 2^23 doubles (64 MiB) that each device the
 * library runs on holds until the process ends; meshes with ne (nm + 2)
 * above that are rejected (SPOC_EINVAL).  No NaN test (the values are not
 * read on the host).  All calls on one device share that array: calls that
 * overlap on one device (several contexts on several host threads or
 * streams) race; calls queued on one context's stream do not.
 */
int spoc_energy_noalloc(spoc_ctx *ctx, double *out_dev);

/*
 * Duration (milliseconds) of the last spoc_rk4_steps call measured with HIP
 * events on the context stream, and the per-kernel totals of that call:
 * names[i] (NUL-separated, caller buffer of len bytes), ms[i], launches[i]
 * for i < *nkernels (in/out: capacity in, count out).
 */
int spoc_last_timing(spoc_ctx *ctx, double *total_ms, int *nkernels, char *names, int len,
                     double *ms, int *launches);
int spoc_set_profiling(spoc_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
