/*
 * viennals_gpu.h -- C ABI of the MI355X-native level-set advection library
 * (libviennals_gpu.so).  Plain pointers and sizes only.
 *
 * Each entry point replaces one operation of the reference ViennaLS v5.8.3
 * (header-only C++, include/viennals/...); the reference interface is cited
 * beside each declaration.  All functions return 0 on success and a negative
 * code on error; lsg_last_error() returns the message of the last error of
 * the calling thread (the role of ViennaCore's Logger error in the
 * reference, e.g. lsAdvect.hpp:1019-1027).
 *
 * Level sets live in GPU memory.  A level set is exchanged with the host as
 * its run-start list in HRLE iteration order (x fastest): every defined grid
 * point with its value, and every undefined grid point that starts a run of
 * a new sign, with value +DBL_MAX or -DBL_MAX (Domain::POS_VALUE/NEG_VALUE,
 * lsDomain.hpp:51-52).  Entry 0 is the grid minimum point.
 */
#ifndef VIENNALS_GPU_H
#define VIENNALS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lsg_domain lsg_domain;
typedef struct lsg_advect lsg_advect;

/* Boundary conditions: viennahrle::BoundaryType numbering
 * (python/viennals/_core.pyi:26-30). */
enum { LSG_REFLECTIVE = 0, LSG_INFINITE = 1, LSG_PERIODIC = 2, LSG_POS_INFINITE = 3, LSG_NEG_INFINITE = 4 };

/* Spatial schemes: SpatialSchemeEnum (lsAdvectIntegrationSchemes.hpp:10-23). */
enum {
  LSG_ENGQUIST_OSHER_1ST_ORDER = 0,
  LSG_ENGQUIST_OSHER_2ND_ORDER = 1,
  LSG_LAX_FRIEDRICHS_1ST_ORDER = 2,
  LSG_LAX_FRIEDRICHS_2ND_ORDER = 3,
  LSG_LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER = 4,
  LSG_LOCAL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 5,
  LSG_LOCAL_LOCAL_LAX_FRIEDRICHS_2ND_ORDER = 6,
  LSG_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 7,
  LSG_LOCAL_LAX_FRIEDRICHS_2ND_ORDER = 8,
  LSG_STENCIL_LOCAL_LAX_FRIEDRICHS_1ST_ORDER = 9,
  LSG_WENO_3RD_ORDER = 10,
  LSG_WENO_5TH_ORDER = 11,
  /* No counterpart in the reference: WENO_5TH_ORDER whose nonlinear weights
   * come from a trained network (lsg_advect_set_learned_weights) wherever the
   * smoothness indicators agree within a factor 10; plain weno5 elsewhere. */
  LSG_WENO_5TH_ORDER_LEARNED = 12
};

/* Temporal schemes: TemporalSchemeEnum (lsAdvectIntegrationSchemes.hpp:31-35). */
enum { LSG_FORWARD_EULER = 0, LSG_RUNGE_KUTTA_2ND_ORDER = 1, LSG_RUNGE_KUTTA_3RD_ORDER = 2 };

/* BooleanOperationEnum (lsBooleanOperation.hpp:27-33). */
enum { LSG_INTERSECT = 0, LSG_UNION = 1, LSG_RELATIVE_COMPLEMENT = 2 };

/* Built-in velocity functors (evaluated on the GPU), replacing the virtual
 * VelocityField<T> (lsVelocityField.hpp:11-41):
 *   getScalarVelocity = b                                   (LSG_VEL_CONSTANT)
 *                     = b + (n[dir] > 0 ? aPos : aNeg) * |n[dir]|^k
 *                                                           (LSG_VEL_NORMAL_POWER)
 *                     = b + aPos * (|n[0]| + |n[1]| + |n[2]|) (LSG_VEL_NORMAL_SUM,
 *                       the velocity of examples/AirGapDeposition/AirGapDeposition.cpp:20-27)
 *   getVectorVelocity = (vec[0], vec[1], vec[2])            (every kind) */
enum { LSG_VEL_CONSTANT = 0, LSG_VEL_NORMAL_POWER = 1, LSG_VEL_NORMAL_SUM = 2 };

const char *lsg_last_error(void);

/* ---- Domain<T,D> (lsDomain.hpp:27-269) ---------------------------------- */

/* Domain(bounds, boundaryConditions, gridDelta) (lsDomain.hpp:68-78); gmin/gmax
 * are the index bounds floor(b0/dx), ceil(b1/dx) per dimension. D is 2 or 3. */
int lsg_domain_create(int D, const int32_t *gmin, const int32_t *gmax, const int *bc, double grid_delta,
                      lsg_domain **out);
int lsg_domain_destroy(lsg_domain *d);
/* Domain::deepCopy (lsDomain.hpp:124-129). */
int lsg_domain_deep_copy(lsg_domain *dst, const lsg_domain *src);
/* Domain::insertPoints (lsDomain.hpp:135-138): n defined points idx[n*D], val[n]. */
int lsg_domain_insert_points(lsg_domain *d, int64_t n, const int32_t *idx, const double *val, int width);
/* Set / get the run-start list (see top of file). */
int lsg_domain_set_runs(lsg_domain *d, int64_t n, const int32_t *idx, const double *val, int width);
int lsg_domain_num_runs(lsg_domain *d, int64_t *n);
int lsg_domain_get_runs(lsg_domain *d, int32_t *idx, double *val, int64_t cap);
/* Domain::getNumberOfPoints (lsDomain.hpp:156), getLevelSetWidth (:158). */
int lsg_domain_num_points(lsg_domain *d, int64_t *n);
int lsg_domain_width(lsg_domain *d, int *width);
int lsg_domain_set_width(lsg_domain *d, int width); /* Domain::setLevelSetWidth (lsDomain.hpp:160) */

/* Domain::getPointData (lsDomain.hpp:44, 166-168; PointData of ViennaCore):
 * named arrays of `components` = 1 (scalar) or 3 (vector) doubles per
 * defined point, in point-id order, kept on the device and translated by
 * every band operation (the newDataSourceIds of Expand, Reduce, Prune,
 * BooleanOperation and Advect::rebuildLS).  insert: append
 * (insertNextScalarData / insertNextVectorData, replace = 0) or replace the
 * first array of that label (insertReplaceScalarData, replace = 1); n tuples.
 * count / label / size / get address the arrays of one kind by index in
 * insertion order (label returns the label's length, negative on error);
 * size is the tuple count after the last translation (every array holds
 * one tuple per recorded source then). */
int lsg_domain_point_data_insert(lsg_domain *d, const char *label, int components, const double *values, int64_t n,
                                 int replace);
int lsg_domain_point_data_count(const lsg_domain *d, int components);
int lsg_domain_point_data_label(const lsg_domain *d, int components, int i, char *buf, int len);
int64_t lsg_domain_point_data_size(const lsg_domain *d, int components, int i);
int lsg_domain_point_data_get(const lsg_domain *d, int components, int i, double *out, int64_t n);
int lsg_domain_point_data_clear(lsg_domain *d);
/* Domain::serialize (lsDomain.hpp:188-215), the bytes Writer puts in a
 * .lvst file (lsWriter.hpp:35-60): "lsDomain", format version 0, the grid,
 * the run list, the width as uint32 and a point-data flag byte with the
 * point data after it.  The grid, run-list and point-data sections are
 * this library's layout (the reference takes them from ViennaHRLE and
 * ViennaCore, which it does not contain).  *size receives the byte count;
 * with buf null nothing is written, else at most cap bytes. */
int lsg_domain_serialize(lsg_domain *d, uint8_t *buf, int64_t cap, int64_t *size);
/* Domain::deserialize (lsDomain.hpp:218-268), Reader (lsReader.hpp): the
 * grid, band, width and (when the stream has any) point data are replaced;
 * a missing header (-60), a newer version (-62), another dimension (-63)
 * or a short or malformed stream (-61) leave the level set unchanged. */
int lsg_domain_deserialize(lsg_domain *d, const uint8_t *buf, int64_t size);
/* Domain::getGrid (lsDomain.hpp:141-144): dimension, index bounds as given,
 * boundary conditions (BoundaryConditionEnum values) and grid delta. */
int lsg_domain_get_grid(const lsg_domain *d, int *D, int32_t *gmin, int32_t *gmax, int *bc, double *dx);
/* MarkVoidPoints<T,D>::apply with its default options, as Advect uses it
 * (lsMarkVoidPoints.hpp:156-331, lsAdvect.hpp:443-447): point data
 * "VoidPointMarkers", 1 for a point enclosed away from the top region. */
int lsg_mark_void_points(lsg_domain *d);

/* Surface points of a level set: the zero crossing on every grid edge between
 * face-adjacent defined points of opposite sign, linearly interpolated, in
 * physical coordinates (D doubles each) -- the node positions of
 * ToSurfaceMesh (lsToSurfaceMesh.hpp) without its node merging.  *n receives
 * the count; out may be null to query it. */
int lsg_surface_points(lsg_domain *d, double *out, int64_t cap, int64_t *n);
/* CompareChamfer::apply (lsCompareChamfer.hpp:104-245) on those surface
 * points: out[0..4] = forward (target -> sample), backward, Chamfer, RMS and
 * maximum nearest-neighbour distance; counts[0..1] = target / sample points. */
int lsg_compare_chamfer(lsg_domain *target, lsg_domain *sample, double *out, int64_t *counts);
/* CompareVolume::apply (lsCompareVolume.hpp:176-289; getCellCount,
 * getVolumeMismatch): the cells whose corner average is inside (<= 0) in
 * exactly one of the two level sets, after expanding narrower level sets to
 * width 3 on copies; both must share grid bounds, boundary conditions and
 * gridDelta.  cells and volume may be NULL. */
int lsg_compare_volume(lsg_domain *target, lsg_domain *sample, uint64_t *cells, double *volume);

/* ---- MakeGeometry (lsMakeGeometry.hpp) --------------------------------- */
/* makeSphere (lsMakeGeometry.hpp:177-241). */
int lsg_make_sphere(lsg_domain *d, const double *origin, double radius);
/* MakeGeometry::setIgnoreBoundaryConditions (lsMakeGeometry.hpp:125-141)
 * with a sphere: bit d of ignore_mask leaves the indices of dimension d
 * unmapped when it is periodic (:227-235); points that then lie outside the
 * grid are dropped, as on every insertion. */
int lsg_make_sphere_ignore_bc(lsg_domain *d, const double *origin, double radius, int ignore_mask);
/* makePlane (lsMakeGeometry.hpp:245-360): the plane's mesh over the grid
 * (one line or two triangles reaching one cell beyond the grid bounds along
 * the axes other than the single infinite one) converted by FromSurfaceMesh
 * as for makeBox; boundary elements are always dropped. */
int lsg_make_plane(lsg_domain *d, const double *origin, const double *normal);
/* makeBox (lsMakeGeometry.hpp:363-419): the box's surface mesh (four lines
 * or twelve triangles) converted by FromSurfaceMesh
 * (lsFromSurfaceMesh.hpp:243-480): grid-line crossings, distances along the
 * lines times 1 - 1e-7, clipped to +-1, the smallest sign distance per point;
 * mesh elements on the grid boundary are dropped except along periodic axes.
 * The _ignore_bc variant takes setIgnoreBoundaryConditions per axis (bit d):
 * then boundary elements are dropped along periodic axis d too. */
int lsg_make_box(lsg_domain *d, const double *min_corner, const double *max_corner);
int lsg_make_box_ignore_bc(lsg_domain *d, const double *min_corner, const double *max_corner, int ignore_mask);

/* ---- band maintenance ----------------------------------------------------- */
/* Expand<T,D>(ls, width).apply() (lsExpand.hpp:44-172). */
int lsg_expand(lsg_domain *d, int width);
/* lsg_expand without allocating device memory.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under graph
 * capture, serving loops, harnesses that forbid it).  The level set's two
 * band stores are the input and the output; the kernels' flags, statistics
 * and tile status words live in a static device array of about 1 MiB per
 * GPU, kept until the process ends.  Limits: a frame of at most 2^22 bitmap
 * words (2^28 grid points), no point data, the expanded band at least 12
 * cells from its frame edge along infinite axes; otherwise an error with
 * the level set unchanged.  When the output exceeds the stores' capacity
 * (-40) the level set holds the cycles completed before, its width says
 * how many.  Calls share the array: calls that overlap on one GPU (from
 * several threads) race.  Synchronises once per cycle. */
int lsg_expand_noalloc(lsg_domain *d, int width);
/* Reduce<T,D>(ls, width).apply() (lsReduce.hpp:55-134). */
int lsg_reduce(lsg_domain *d, int width);
/* Prune<T,D>(ls) with setRemoveStrayZeros (lsPrune.hpp:74-194). */
int lsg_prune(lsg_domain *d, int remove_stray_zeros);
/* BooleanOperation<T,D>(a, b, op).apply() (lsBooleanOperation.hpp:59-213,
 * 329-366) for op INTERSECT (0), UNION (1), RELATIVE_COMPLEMENT (2); the
 * result is pruned as with the default pruneResult.  op INVERT (3) negates a
 * (lsBooleanOperation.hpp:214-247); b may then be null. */
int lsg_boolean(lsg_domain *a, const lsg_domain *b, int op);
/* Advect::rebuildLS (lsAdvect.hpp:238-429) on a single level set. */
int lsg_rebuild(lsg_domain *d);

/* ---- Advect<T,D> (lsAdvect.hpp:54-1044) ---------------------------------- */
int lsg_advect_create(lsg_advect **out);
int lsg_advect_destroy(lsg_advect *a);
/* insertNextLevelSet (lsAdvect.hpp:856), clearLevelSets (:860). */
int lsg_advect_insert_next_level_set(lsg_advect *a, lsg_domain *d);
int lsg_advect_clear_level_sets(lsg_advect *a);
/* setVelocityField (lsAdvect.hpp:864) with a built-in functor. */
int lsg_advect_set_velocity_builtin(lsg_advect *a, int kind, int dir, int k, double b, double a_pos, double a_neg,
                                    const double *vec3);
/* Several level sets (Advect::insertNextLevelSet called more than once,
 * lsAdvect.hpp:856): the reference's VelocityField receives the material
 * index, the number of the level set below or at the point
 * (lsVelocityField.hpp:17-38, lsAdvect.hpp:513-527).  This sets the
 * built-in functor of one material (0..7; materials without one use the
 * functor of lsg_advect_set_velocity_builtin, which clears these). */
int lsg_advect_set_velocity_builtin_material(lsg_advect *a, int material, int kind, int dir, int k, double b,
                                             double a_pos, double a_neg, const double *vec3);
/* VelocityField::getDissipationAlpha (lsVelocityField.hpp:32-38) of the
 * built-in functors, used by LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER
 * (lsLocalLaxFriedrichsAnalytical.hpp): a constant per direction, for the
 * default functor (material -1) or one material's; the velocity setters
 * above reset it to 0 (the reference's default). */
int lsg_advect_set_dissipation_alpha_builtin(lsg_advect *a, int material, const double *alpha3);
/* setVelocityField (lsAdvect.hpp:864) with a VelocityField<T> subclass
 * evaluated on the host (lsVelocityField.hpp:17-38): for every active point
 * of every stage the library calls scalar(coordinate, material, normal,
 * pointId, user) and vector(..., out[3]) with coordinate = index * gridDelta
 * (3 components), material 0, the normal the scheme computes and pointId =
 * rank among the defined points of the expanded band, in point order, on the
 * calling thread.  Either callback may be null (zero).  Supported with the
 * Engquist-Osher, Lax-Friedrichs (also findGlobalAlphas' calls, lsAdvect.hpp:
 * 90-195), local Lax-Friedrichs (also the 3^D alpha box, lsLocalLaxFriedrichs
 * .hpp:184-232), analytical local Lax-Friedrichs (with
 * lsg_advect_set_dissipation_alpha_host), local-local Lax-Friedrichs, WENO
 * and stencil local Lax-Friedrichs schemes (the point and its 3^D box, then
 * scalar() at every box normal moved by -DN and +DN per axis,
 * lsStencilLocalLaxFriedrichsScalar.hpp:252-338); not with the learned
 * WENO5 scheme. */
typedef double (*lsg_scalar_velocity_fn)(const double *coordinate, int material, const double *normal,
                                         uint64_t point_id, void *user);
typedef void (*lsg_vector_velocity_fn)(const double *coordinate, int material, const double *normal,
                                       uint64_t point_id, void *user, double *out3);
int lsg_advect_set_velocity_host(lsg_advect *a, lsg_scalar_velocity_fn scalar, lsg_vector_velocity_fn vector,
                                 void *user);
/* VelocityField::getDissipationAlpha (lsVelocityField.hpp:32-38) of a
 * host-evaluated velocity field, for LOCAL_LAX_FRIEDRICHS_ANALYTICAL_1ST_ORDER
 * (lsLocalLaxFriedrichsAnalytical.hpp:186-214): called for every direction
 * at each of the 3^D box points of every active point, with the box point's
 * central differences (not normalised), after the velocity callbacks of
 * lsg_advect_set_velocity_host; alpha(direction) is the maximum over the box.
 * fn = NULL gives the reference's default, 0.  A built-in velocity setter
 * removes it. */
typedef double (*lsg_dissipation_alpha_fn)(int direction, int material, const double *central_differences,
                                           void *user);
int lsg_advect_set_dissipation_alpha_host(lsg_advect *a, lsg_dissipation_alpha_fn fn, void *user);
int lsg_advect_set_spatial_scheme(lsg_advect *a, int scheme);     /* :936 */
int lsg_advect_set_temporal_scheme(lsg_advect *a, int scheme);    /* :949 */
int lsg_advect_set_advection_time(lsg_advect *a, double time);    /* :873 */
int lsg_advect_set_time_step_ratio(lsg_advect *a, double cfl);    /* :885 */
int lsg_advect_set_single_step(lsg_advect *a, int single);        /* :879 */
int lsg_advect_set_dissipation_alpha(lsg_advect *a, double alpha);/* :955 */
int lsg_advect_set_check_dissipation(lsg_advect *a, int check);   /* :958 */
int lsg_advect_set_calculate_normal_vectors(lsg_advect *a, int c);/* :891 */
int lsg_advect_set_ignore_voids(lsg_advect *a, int iv);           /* :899 */
int lsg_advect_set_adaptive_time_stepping(lsg_advect *a, int on, unsigned subdivisions); /* :904; acts with several level sets */
int lsg_advect_set_update_point_data(lsg_advect *a, int update);  /* :962; rebuildLS translates or drops point data */
/* Advect::setVelocityUpdateCallback (lsAdvect.hpp:82,966-969), the
 * std::function<bool(SmartPointer<Domain>)>: fn(top, user) runs after the
 * first stage of RK2 and after the first and second stages of RK3, with the
 * advected (top) level set (lsAdvectIntegrationSchemes.hpp:87-89,128-130,
 * 148-150); its return value is ignored, as in the reference, unless it is
 * negative: a callback that failed (an exception in a language binding) ends
 * the apply at that stage with an error, as the exception ends the
 * reference's apply (the level set keeps the stages done).  With several
 * level sets a callback also makes the first stage adjust the lower level
 * sets and every combine that etched restore them (:72-77,98-103,139-144).
 * The callback may read level sets and change this Advect's velocity field;
 * changing the advected band, or advecting from inside it, returns an error.
 * A step that outgrows its band buffers restarts, so the callback can run
 * again for that step.  fn = NULL removes the callback. */
typedef int (*lsg_velocity_update_fn)(lsg_domain *top, void *user);
int lsg_advect_set_velocity_update_callback(lsg_advect *a, lsg_velocity_update_fn fn, void *user);
/* Advect::setSaveAdvectionVelocities (lsAdvect.hpp:916): point data
 * "AdvectionVelocities" and "Dissipation" after every stage (:762-796). */
int lsg_advect_set_save_advection_velocities(lsg_advect *a, int save);
/* Advect::prepareLS (lsAdvect.hpp:973-1015). */
int lsg_advect_prepare_ls(lsg_advect *a);
/* Advect::apply (lsAdvect.hpp:1017-1043). */
int lsg_advect_apply(lsg_advect *a);
int lsg_advect_get_advected_time(lsg_advect *a, double *t);       /* :920 */
int lsg_advect_get_current_time_step(lsg_advect *a, double *dt);  /* :923 */
int lsg_advect_get_number_of_time_steps(lsg_advect *a, unsigned *n); /* :926 */
/* No counterpart in the reference: active-point updates of the last apply()
 * (sum over stages of points with |phi| <= 0.5) and, when profiling is on,
 * the summed duration (ms), launch count and algorithmic bytes
 * (8*band entries + 8*active points per launch) of the rate kernel. */
int lsg_advect_get_point_updates(lsg_advect *a, double *n);
int lsg_advect_set_profiling(lsg_advect *a, int on);
int lsg_advect_get_rate_kernel_stats(lsg_advect *a, double *ms, int64_t *launches, double *bytes);
/* No counterpart in the reference: the same figures for the evaluation
 * kernel of the star operations (Expand cycles and rebuilds), the kernel
 * that takes most of a step; algorithmic bytes per launch are 8 per input
 * band entry plus 8 per candidate. */
int lsg_advect_get_star_kernel_stats(lsg_advect *a, double *ms, int64_t *launches, double *bytes);
/* No counterpart in the reference: the same figures for the evaluation
 * kernel of two Expand cycles run as one band operation (the cycles of an
 * Expand go two at a time on one level set without point data); bytes
 * counted as above. */
int lsg_advect_get_expand2_kernel_stats(lsg_advect *a, double *ms, int64_t *launches, double *bytes);

/* Finite-difference kernels FiniteDifferences::weno3/weno5
 * (lsFiniteDifferences.hpp:58-123) evaluated on the GPU for n stencils:
 * x[n*(2r+1)] -> out[n]. */
int lsg_weno_eval(int order, int64_t n, const double *x, double dx, int plus, double *out);

/* ---- Learned WENO5 weights (no counterpart in the reference) -------------
 * A network 5 -> 32 -> 32 -> 3 (GELU, softmax) replaces the nonlinear
 * weights alpha_k / sum(alpha) of FiniteDifferences::weno5
 * (lsFiniteDifferences.hpp:108-122).  Parameters: 1347 floats laid out
 * W1[32][5], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]; the GPU uses
 * them as bf16 with fp32 accumulation. */
typedef struct lsg_mlp lsg_mlp;
/* Network parameters of spatial scheme LSG_WENO_5TH_ORDER_LEARNED. */
int lsg_advect_set_learned_weights(lsg_advect *a, const float *params, int n);
/* Learned weno5 on n stencils x[n*7] (same arguments as lsg_weno_eval). */
int lsg_weno5_learned_eval(int64_t n, const double *x, double dx, int plus, const float *params, double *out);
/* Trainer: fp32 master parameters, Adam (beta 0.9/0.999, eps 1e-8).  A sample
 * is 8 floats (the 5 inputs d/max|d| and the 3 candidate polynomials / max|d|)
 * and a target (derivative * dx / max|d|); loss = mean squared error of
 * sum_k w_k p_k. */
int lsg_mlp_create(lsg_mlp **out, const float *params);
int lsg_mlp_destroy(lsg_mlp *m);
int lsg_mlp_set_samples(lsg_mlp *m, int64_t n, const float *x8, const float *target);
/* Loss and gradient over samples [offset, offset + count). */
int lsg_mlp_loss_grad(lsg_mlp *m, int64_t offset, int64_t count, double *loss, float *grad);
/* iters Adam steps over consecutive minibatches of `batch` samples. */
int lsg_mlp_train(lsg_mlp *m, int iters, int64_t batch, double lr, double *last_loss);
int lsg_mlp_get_params(lsg_mlp *m, float *out);

/* ---- Batches of independent level sets (no counterpart in the reference) ---
 * Many Advect<T,D> instances (lsAdvect.hpp:54-1044), one level set each,
 * stepped together: the bands are concatenated into one segmented band on
 * the device, every kernel of a step is one launch for all level sets, the
 * step is captured into a HIP graph, and each level set keeps its own grid,
 * velocity and CFL time step.  lsg_batch_apply is Advect::apply with
 * advectionTime 0 (one step) for every level set.  Schemes: Engquist-Osher,
 * WENO and the Lax-Friedrichs schemes LF, LLF and LLLF (alphas and CFL
 * bound per level set).  Results are written back into the domains by
 * lsg_batch_write_back. */
typedef struct lsg_batch lsg_batch;
int lsg_batch_create(lsg_batch **out, lsg_domain *const *domains, int n);
int lsg_batch_destroy(lsg_batch *b);
int lsg_batch_set_velocity_builtin(lsg_batch *b, int k, int kind, int dir, int kk, double b0, double a_pos,
                                   double a_neg, const double *vec3);
int lsg_batch_set_schemes(lsg_batch *b, int spatial, int temporal, double cfl);
int lsg_batch_apply(lsg_batch *b);
/* Accumulated advected time of every level set (n values). */
int lsg_batch_get_advected_times(lsg_batch *b, double *t);
int lsg_batch_get_point_updates(lsg_batch *b, double *n, unsigned *steps);
int lsg_batch_write_back(lsg_batch *b);
/* Grid of the scan kernels (blocks), or chosen by timing a step per
 * candidate grid (the second of two from the same band; the band is
 * restored afterwards). */
int lsg_batch_set_scan_blocks(lsg_batch *b, int blocks);
int lsg_batch_autotune(lsg_batch *b, int *chosen, double *ms);

/* Device synchronisation (for timing by callers). */
int lsg_device_synchronize(void);
/* Band operations of single level sets and batches as one single-pass
 * kernel each (on != 0) or as the multi-kernel pipeline (0, the default;
 * also set by LSG_FUSED_BANDOPS=1).  A tuning control with no counterpart
 * in the reference; both give the same level sets.  Advect objects keep
 * the kernels of the steps they have captured. */
int lsg_set_single_pass_band_ops(int on);

#ifdef __cplusplus
}
#endif

#endif
