/*
 * tah.h -- C ABI of the MI355X ghost-fill engine for TreeAMR.jl.
 *
 * The reference has no FFI: its fill is the Julia call
 *     fill_ghosts!(fs::FieldSet, schedule::GhostSchedule; boundary=nothing)
 * (src/ghosts.jl:1783-1794), which checks the pair (check_fill,
 * src/ghosts.jl:2058-2115) and replays the schedule stage by stage
 * (exchange_ghosts!, src/ghosts.jl:1099-1160), one KernelAbstractions
 * launch per transfer group (run_group!, src/ghosts.jl:420-499).
 * This header is what a Julia shim `ccall`s in its place: the schedule is
 * flattened once per (schedule, forest generation) into a tah_schedule_desc
 * and uploaded as a tah_plan; every fill then runs on the device.
 *
 * Conventions: every function returns a tah_status; on failure
 * tah_last_error() holds a message (thread-local). Nothing is launched when
 * a check fails. Indices in descriptors are 0-based (the reference's are
 * 1-based: a shim subtracts one). Field data is Float64 only
 * (check_floattype, src/device.jl:53-61). Device pointers are hipMalloc'd
 * by the caller, sized by tah_layout_size; streams are hipStream_t passed as
 * void*.
 * Two fills over one working array must not overlap (src/ghosts.jl:1773).
 */
#ifndef TAH_H
#define TAH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tah_status {
  TAH_OK = 0,
  TAH_EINVAL = 1,      /* bad argument (ArgumentError in the reference) */
  TAH_ESTALE = 2,      /* schedule built for another forest generation */
  TAH_ELAYOUT = 3,     /* work layout does not match the schedule */
  TAH_EHIP = 4,        /* a HIP runtime call failed */
  TAH_ENOMEM = 5,      /* device allocation failed */
  TAH_EUNSUPPORTED = 6 /* a feature this library does not implement */
} tah_status;

/* Message of the last failure on this thread. */
const char* tah_last_error(void);

/* ------------------------------------------------------------------ */
/* Host side: the forest, the operators and the schedule builder.       */
/* ------------------------------------------------------------------ */

/* unit_lagrange_weights(p, num//den) (src/operators.jl:675-682): the
 * weights of nodes 0..p-1 at num/den, exact and rounded once to double. */
tah_status tah_lagrange_weights(int32_t p, int64_t num, int64_t den, double* out);
/* conservative_prolong_weights(p) (src/schedule.jl:1097-1120): the low and
 * high subcell weight vectors of odd order p. */
tah_status tah_conservative_weights(int32_t p, double* low, double* high);

typedef struct tah_forest tah_forest;
typedef struct tah_schedule tah_schedule;

/* Forest(roots; N, periodic, reflecting, outflow, extents)
 * (src/forest.jl:314-341). reflecting/outflow: [dim][2] flags (lo, hi);
 * extents: [dim][2] or NULL for (0, roots[d]). */
tah_status tah_forest_create(int32_t dim, const int32_t* roots, int32_t N,
                             const int32_t* periodic, const int32_t* reflecting,
                             const int32_t* outflow, const double* extents,
                             tah_forest** out);
tah_status tah_forest_destroy(tah_forest* f);
/* nleaves (src/forest.jl:1052) and generation (src/forest.jl:1040). */
tah_status tah_forest_info(const tah_forest* f, int32_t* nleaves, int64_t* generation);
/* The sorted leaf keys: root, level, coords[3] per leaf (src/morton.jl:25). */
tah_status tah_forest_leaves(const tah_forest* f, int32_t* root, int32_t* level,
                             uint32_t* coords);
/* refine! of the given leaves (0-based leaf indices), src/forest.jl:1761. */
tah_status tah_forest_refine(tah_forest* f, int32_t n, const int32_t* leaves);
/* balance! (src/forest.jl:1837) and isbalanced (src/forest.jl:1892). */
tah_status tah_forest_balance(tah_forest* f);
tah_status tah_forest_isbalanced(const tah_forest* f, int32_t* balanced);
/* Coarsening, as regrid!'s coarsening flags (src/regrid.jl): every family
 * whose 2^dim children are all among the given leaves (0-based indices)
 * becomes its parent, unless the parent would meet a leaf two levels finer
 * (2:1 balance holds after). coarsened (may be NULL): families merged. */
tah_status tah_forest_coarsen(tah_forest* f, int32_t n, const int32_t* leaves, int32_t* coarsened);
/* Positions of the stored points along dimension d (0-based) of every leaf:
 * out[b*S + j] for stored index j = 0 .. S-1 of leaf b, ghosts included, of a
 * set with ghost width G and centering c (0 cell-like, 1 vertex-like) along
 * d. Formed as nodeposition does (src/storage.jl:1008-1041, with the anchor,
 * integer base and spacing of src/geometry.jl:57,76-131): the integer k =
 * B + j - G counts spacings h of the leaf's level from the anchor lo (zero
 * where zero is a node of the level-0 grid, else the low corner), and the
 * position is fma(k, h, lo) (vertex-like) or fma(2k+1, h/2, lo) (cell-like),
 * rounded once; so every copy of a point is at one position bit for bit. */
tah_status tah_forest_positions(const tah_forest* f, int32_t d, int32_t G, int32_t c, int32_t S,
                                double* out);

enum { TAH_POINTVALUE = 0, TAH_CONSERVATIVE = 1 };
/* Operators(; prolongation, restriction, family, bounded,
 * vertex_prolongation) (src/operators.jl:314-415). */
typedef struct tah_operators {
  int32_t family;
  int32_t prolongation;
  int32_t restriction;
  int32_t bounded;
  int32_t vertex_prolongation; /* 0 for none */
} tah_operators;
/* The constructor's own checks (src/operators.jl:327-411). */
tah_status tah_operators_check(const tah_operators* ops);

enum { TAH_PARITY_NONE = 0, TAH_PARITY_EVEN = 1, TAH_PARITY_ODD = 2 };
enum { TAH_CELL = 0, TAH_VERTEX = 1 };

/* GhostSchedule(forest, ops; G, centering, parity)
 * (src/schedule.jl:3075-3292). G, centering: per dimension; parity:
 * nvars x dim (variable-major) or NULL. */
tah_status tah_schedule_create(const tah_forest* f, const tah_operators* ops,
                               const int32_t* G, const int32_t* centering,
                               int32_t nvars, const int8_t* parity,
                               tah_schedule** out);
tah_status tah_schedule_destroy(tah_schedule* s);
/* isstale (src/schedule.jl:872-874): TAH_ESTALE once the forest changed. */
tah_status tah_schedule_check(const tah_schedule* s);

/* The factor table of the mirrored transfers, factors[v + nvars*(col-1)],
 * col = 1..3^dim (parityfactors, src/storage.jl:536-550). */
tah_status tah_parity_factors(int32_t dim, int32_t nvars, const int8_t* parity,
                              double* out);

/* --- The flattened schedule (north-star (a)) ------------------------- */

enum { TAH_COPY = 1, TAH_RESTRICT = 2, TAH_PROLONG = 3 };
/* Stage tags, in fill order: stage P (-1), stage 0 (0), phase 1 (1), the
 * prolongations onto level l (2 + l) (transfer_stage,
 * src/schedule.jl:640-658). */
enum { TAH_STAGE_P = -1, TAH_STAGE_UNIFY = 0, TAH_STAGE_PHASE1 = 1 };
/* The regrid transfer's two stages (REGRID_TAG and REGRID_KEPT_TAG,
 * src/regrid.jl:755-803, 905-933): old array -> new array, then, for a set
 * with a vertex-like dimension, the kept blocks' copies of shared points
 * over the created blocks' within the new array. */
enum { TAH_STAGE_REGRID = 100, TAH_STAGE_REGRID_KEPT = 101 };

/* One dimension of a transfer (Stencil1D, src/schedule.jl:64-93):
 * target rows targetfirst .. targetfirst+ntarget-1 (0-based); row i reads
 * source rows srcstart[i] .. srcstart[i]+order-1 with weights
 * weights[order*i .. order*i+order-1]. */
typedef struct tah_stencil1d {
  int32_t targetfirst;
  int32_t ntarget;
  int32_t order;
  int32_t unit; /* one point of weight one */
  const int32_t* srcstart;
  const double* weights;
} tah_stencil1d;

/* One transfer group (TransferGroup, src/schedule.jl:150-197). Dimensions
 * beyond dim are one-point unit stencils. */
typedef struct tah_group_desc {
  int32_t kind;       /* TAH_COPY / TAH_RESTRICT / TAH_PROLONG */
  int32_t stage;      /* index into stage_tags */
  int32_t factorcol;  /* 0: unscaled, else the 1-based factor column */
  int32_t ntransfers;
  const int32_t* targetblocks; /* 0-based, non-decreasing */
  const int32_t* sourceblocks; /* 0-based */
  tah_stencil1d stencils[3];
} tah_group_desc;

/* One outer-face hook region (BoundaryRegion, src/schedule.jl:200-229;
 * boundary_region, :2188-2217): block, direction, and the 0-based half-open
 * box lo..hi of stored indices the hook writes. For the relative hook:
 * clamps[d], the stored index of the nearest evolved point along a
 * component of the direction that leaves the domain through an outer face
 * (-1 along every other, where the point keeps the ghost's own index), and
 * slot, the tag of the stage after which the ghosts that point reads are
 * filled (TAH_STAGE_PHASE1, or the prolongations onto the block's level).
 * first: the index of the region's first point among every region's
 * points, a region's points numbered with i fastest, then j, then k. */
typedef struct tah_boundary_region {
  int32_t block;
  int32_t direction[3];
  int32_t lo[3];
  int32_t hi[3];
  int32_t clamps[3];
  int32_t slot;
  int64_t first;
} tah_boundary_region;

typedef struct tah_schedule_desc {
  int32_t dim;
  int32_t nblocks;
  int32_t stored[3];       /* N + 2G + c per dimension; 1 beyond dim */
  int32_t nstages;
  const int32_t* stage_tags;
  /* stagegroups (src/schedule.jl:2876-2931): stages [first[k], last[k]]
   * may run as one. */
  int32_t nstagegroups;
  const int32_t* stagegroup_first;
  const int32_t* stagegroup_last;
  int32_t ngroups;
  const tah_group_desc* groups; /* sorted by stage */
  int64_t generation;           /* forest generation it was built for */
  int32_t nsources;             /* regrid: the old leaves the sources index; 0: nblocks */
  int32_t pad_;
  /* The outer-face hook regions, in tah_schedule_boundaries' order, and
   * their points in all (none for a regrid or unify descriptor). */
  int32_t nboundaries;
  int32_t pad2_;
  const tah_boundary_region* boundaries;
  int64_t boundary_points;
} tah_schedule_desc;

/* The regrid transfer of one field-set layout onto the forest's current
 * leaves from its leaves before the regrid (old_root, old_level,
 * old_coords[3] per old leaf, sorted along the curve, as tah_forest_leaves
 * gave them): transfer_groups / regrid_stage (src/regrid.jl:755-803,
 * 935-947) -- each new leaf from the same old leaf (a copy of its own
 * range), its old parent (a prolongation) or its old children
 * (restrictions), the ghost builders' stencils at direction zero -- and,
 * for a set with a vertex-like dimension, the kept-over-created pass
 * (regrid_kept_pairs / regrid_kept_stage, src/regrid.jl:826-933). Stage
 * tags TAH_STAGE_REGRID (sources index the old leaves, desc.nsources of
 * them) and TAH_STAGE_REGRID_KEPT (sources index the new ones). A plan of
 * it (tah_plan_create, with the new array's layout) runs with
 * tah_regrid_transfer only. */
tah_status tah_regrid_schedule_create(const tah_forest* f, int32_t nold, const int32_t* old_root,
                                      const int32_t* old_level, const uint32_t* old_coords,
                                      const tah_operators* ops, const int32_t* G,
                                      const int32_t* centering, int32_t nvars, const int8_t* parity,
                                      tah_schedule** out);

/* UnifySchedule(fs) (src/unify.jl:76-170): stage P and stage 0 alone, for a
 * set with a vertex-like dimension (refused otherwise), with the
 * restriction the geometry fixes -- injection along a vertex-like
 * dimension, the exact two-cell average along a cell-like one -- and any
 * ghost width G >= 0 (the set needs no ghosts). */
tah_status tah_unify_schedule_create(const tah_forest* f, const int32_t* G, const int32_t* centering,
                                     int32_t nvars, const int8_t* parity, tah_schedule** out);
/* What unify! replays (unify_arrays!, src/ghosts.jl:1874-1906): stages P and
 * 0 of a ghost or unify schedule as a schedule of their own, in the working
 * array's indices (state = 0) or a state vector's (state = 1: every index
 * shifted by -G, stored = the own extents, for a dense layout of the state
 * vector). A plan of it filled with tah_fill is unify!. Valid while s
 * lives. */
tah_status tah_unify_desc_get(tah_schedule* s, int32_t state, const tah_schedule_desc** out);

/* The schedule's flattened form; pointers stay valid while s lives. */
tah_status tah_schedule_desc_get(const tah_schedule* s, const tah_schedule_desc** out);

/* The outer-face regions (the reference's hook regions, filled by the
 * application, never by the fill): n, then per region block, direction[3]
 * and the 0-based half-open box lo[3], hi[3]. */
tah_status tah_schedule_boundaries(const tah_schedule* s, int32_t* n, int32_t* block,
                                   int32_t* direction, int32_t* lo, int32_t* hi);

/* Effective values written per fill (bench/copies.jl:115-125): the sum
 * over groups of ntransfers x box points, times nvars by the caller. */
tah_status tah_schedule_points(const tah_schedule* s, int64_t* points);

/* ------------------------------------------------------------------ */
/* Device side.                                                          */
/* ------------------------------------------------------------------ */

/* The working array: element (i, j, k, v, b) lives at
 * stride[0]*i + stride[1]*j + stride[2]*k + stride[3]*v + stride[4]*b.
 * The reference's dense layout work[i,j,k,var,block] (src/storage.jl:196-200)
 * is strides (1, S0, S0*S1, S0*S1*S2, nvars*S0*S1*S2). Any strides under
 * which blocks are outermost and no two elements meet are accepted. */
typedef struct tah_layout {
  int32_t stored[3];
  int32_t nvars;
  int32_t nblocks;
  int32_t pad_;
  int64_t stride[5];
} tah_layout;

/* The dense layout of a schedule's field set. */
tah_status tah_layout_dense(const tah_schedule_desc* d, int32_t nvars, tah_layout* out);
/* The arena layout: variables innermost, strides (nvars, nvars*S0,
 * nvars*S0*S1, 1, block) with every block starting on a 256-byte boundary,
 * so that a ghost slab of any orientation is read and written in runs of at
 * least nvars consecutive doubles. */
tah_status tah_layout_arena(const tah_schedule_desc* d, int32_t nvars, tah_layout* out);
/* Elements a layout spans (what to allocate, in doubles). */
tah_status tah_layout_size(const tah_layout* l, int64_t* elements);

typedef struct tah_ctx tah_ctx;
typedef struct tah_plan tah_plan;
typedef struct tah_graph tah_graph;

tah_status tah_ctx_create(int32_t device, tah_ctx** out);
tah_status tah_ctx_destroy(tah_ctx* ctx);

/* Upload a schedule's tables for one layout (todevice, src/device.jl:13-33):
 * the library owns the device copy; the caller keeps its host arrays. Every
 * index is validated against the layout here, before any launch. */
tah_status tah_plan_create(tah_ctx* ctx, const tah_schedule_desc* d, const tah_layout* l,
                           tah_plan** out);
tah_status tah_plan_destroy(tah_plan* p);
/* Values one fill writes (per variable: x nvars already applied). */
tah_status tah_plan_values(const tah_plan* p, int64_t* values);

/* The plain per-group fill: one launch per transfer group, a flat index per
 * written value, as the reference's device path launches
 * (launch_transfer!, src/ghosts.jl:507-521). factors: device table
 * nvars x 3^dim, or NULL when no group is mirrored. */
tah_status tah_fill_plain(tah_plan* p, double* work, const double* factors, void* stream);

/* The regrid transfer (run_stages! of regrid!, src/regrid.jl:1337-1350 and
 * 1368-1384): the plan's first stage reads old_work (the set's array
 * before the regrid, old_layout: the same layout per block, nsources
 * blocks) and writes new_work's own ranges; the second stage copies within
 * new_work. new_work's ghosts are not written (the next fill does), and
 * whatever else it holds stays. As in the reference, the old array's
 * ghosts must be current (a fill first): a prolongation reads the parent's
 * ghost layers. Asynchronous on `stream`. */
tah_status tah_regrid_transfer(tah_plan* p, const tah_layout* old_layout, const double* old_work,
                               double* new_work, void* stream);

/* The staged fill: one launch per stage group over a tile list of every
 * group of its stages, class-specialised kernels. Same results, bit for bit. */
tah_status tah_fill(tah_plan* p, double* work, const double* factors, void* stream);

/* fill_ghosts!(fs, schedule; boundary = CellBoundary(...)) (exchange_ghosts!,
 * src/ghosts.jl:1099-1160; the hook kernels :1177-1318): the staged fill
 * with the outer-face hook run between its stage groups, the hook's values
 * computed by the caller at the positions nodeposition gives (the callback
 * itself cannot cross a C boundary). values and fr are device arrays of
 * desc.boundary_points x nvars doubles, the variables of one point
 * contiguous, points in the regions' numbering (tah_boundary_region.first).
 *   TAH_HOOK_VALUES (CellBoundary(g), its AllVariables and WithPatch forms):
 *     after phase 1's stage group, each hook ghost of variable v becomes
 *     values[p*nvars + v]; fr is not read.
 *   TAH_HOOK_RELATIVE (CellBoundary(Relative(f))): values = f(x_g), fr =
 *     f(x_r), x_r the nearest evolved point; after the stage group that
 *     holds the region's slot stage, each ghost becomes
 *     values + (work[r] - fr), in that order.
 * A schedule without hook regions runs as tah_fill. The prescribed wall
 * (CellBoundary(g; wall = :prescribed), stage W) is not provided: the wall
 * plane stays the application's, the reference's default. */
enum { TAH_HOOK_VALUES = 0, TAH_HOOK_RELATIVE = 1 };
tah_status tah_fill_boundary(tah_plan* p, double* work, const double* factors, int32_t kind,
                             const double* values, const double* fr, void* stream);

/* The staged fill with its launches run one after another on `stream`, each
 * timed by HIP events (for measurement: the same kernels and results as
 * tah_fill). Call with cls, values and ms NULL to get the launch count in
 * *n; then with arrays of *n entries: per launch, in fill order, its kernel
 * class, the values it writes and its duration in ms. No counterpart in the
 * reference, which times whole fills (bench/ghosts.jl:69-78). */
tah_status tah_fill_timed(tah_plan* p, double* work, const double* factors, void* stream,
                          int32_t* n, int32_t* cls, int64_t* values, float* ms);
/* The kernel name of a class, as tah_fill_timed reports them. */
tah_status tah_kernel_class_name(int32_t cls, char* buf, int32_t len);

/* Many independent fills (forests, field sets) in one launch per stage
 * position. */
tah_status tah_fill_batched(tah_ctx* ctx, int32_t n, tah_plan* const* plans,
                            double* const* works, const double* const* factors, void* stream);

/* A batched fill captured once in a hipGraph, replayed by tah_graph_launch.
 * The arrays and tables named at creation must stay alive. */
tah_status tah_graph_create(tah_ctx* ctx, int32_t n, tah_plan* const* plans,
                            double* const* works, const double* const* factors,
                            tah_graph** out);
tah_status tah_graph_launch(tah_graph* g, void* stream);
tah_status tah_graph_destroy(tah_graph* g);

/* The adjoint of the fill (the fill is linear): stages in reverse, each
 * source accumulating w x adj[target] over a precomputed inverted schedule
 * in a fixed order (no float atomics), then the targets zeroed. */
tah_status tah_fill_adjoint(tah_plan* p, double* adj, const double* factors, void* stream);
/* tah_fill_adjoint with its launches timed one by one by HIP events on
 * `stream` (for measurement; the same kernels and results). Call with kind,
 * bytes and ms NULL for the launch count in *n; then with arrays of *n
 * entries: per launch, in order, its kind (0, 1, 2: the z, y, x passes of
 * the interpolations' transpose; 3: the gather; 4: the zeroing of
 * interpolated targets), its algorithmic bytes (8 per value read and per
 * value written, each once) and its duration in ms. */
tah_status tah_fill_adjoint_timed(tah_plan* p, double* adj, const double* factors, void* stream, int32_t* n,
                                  int32_t* kind, int64_t* bytes, float* ms);

/* tah_fill_adjoint without any allocation. No counterpart in the reference:
 * this is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The passes stage their values in a buffer compiled into the library, at
 * most 8 Mi doubles (64 MiB) per GPU, held by each GPU the library runs on
 * until the process ends; a plan whose adjoint stages more is refused
 * (TAH_EUNSUPPORTED). The plan's adjoint tables must exist: the plan's first
 * tah_fill_adjoint call builds them at set-up (TAH_EINVAL otherwise). Every
 * call on one GPU shares that GPU's buffer: calls made one after another
 * are ordered by the library, on any streams, but calls that overlap (from
 * several host threads at once, or graph replays on several streams) race. */
tah_status tah_fill_adjoint_noalloc(tah_plan* p, double* adj, const double* factors, void* stream);

/* A plan's launch options (no counterpart in the reference, whose launches
 * are fixed): fork runs a stage group's kernel classes on four forked
 * streams (1) or in order on the fill's stream (0, default); merge runs
 * every interpolation unit class of one order in one launch (1, default) or
 * each in its own (0); defer moves phase-1 transfers that no later stage
 * reads or overwrites into the last stage group (1, default); fuse is
 * refused with TAH_EUNSUPPORTED unless 0; copy_wgs > 0 launches the copies
 * from a bounded grid, leaving wave slots to interpolations forked beside
 * them; interp_pairs gives each thread of a merged interpolation two
 * neighbouring variables (layouts with variables innermost and an even
 * count; 16-byte loads where the array is 16-byte aligned) or one (0);
 * copy_order orders a copy launch's tiles by source block (0, default: the
 * slabs that share source points are read together) or by target block (1:
 * the writes into one block's ghost layer meet; faster where blocks are
 * small). Setting merge, defer, copy_per_thread, copy_wgs, interp_pairs or
 * copy_order rebuilds the
 * plan's tiles (it waits for the device first);
 * batches and graphs made earlier keep the options they were made with. */
typedef struct tah_plan_options {
  int32_t fork;
  int32_t merge;
  int32_t defer;
  int32_t copy_per_thread; /* the copy kernel's tile: 4, 8 (default) or 16 values per thread */
  int32_t fuse;            /* must be 0: the fused copy-and-interpolation launch is switched off */
  int32_t copy_wgs;        /* copies from a bounded grid of this many workgroups (a multiple of 8), or 0 */
  int32_t interp_pairs;    /* merged interpolations over variable pairs where the layout allows (1, default) */
  int32_t lds;             /* interpolation from LDS copies of its source boxes: bit 2 restriction and bit 3
                              prolongation streamed along z through LDS, every layout (default 4 where the
                              layout holds variables innermost, 12 otherwise); else, where the layout holds
                              variables innermost in even number, bit 0 restriction and bit 1
                              prolongation's unit classes of 64 units or more chunk by chunk; unit kernels
                              for the rest */
  int32_t queue;           /* in a batched fill of plans that all set it: one persistent launch per stage
                              position over a work queue of every plan's tiles (1), or a launch per kernel
                              class (0, default) */
  int32_t copy_order;      /* copy tiles by source block (0, default) or by target block (1) */
  int32_t reserved_[2];    /* must be zero */
} tah_plan_options;
tah_status tah_plan_get_options(const tah_plan* p, tah_plan_options* out);
/* Setting the plan-wide options sets every interpolation family's below to
 * them again. */
tah_status tah_plan_set_options(tah_plan* p, const tah_plan_options* o);

/* The options of one interpolation family: the transfers of order `order`
 * in every dimension that restrict (restriction = 1) or prolongate (0). No
 * counterpart in the reference, whose one kernel serves every transfer
 * (src/ghosts.jl:372-402). */
typedef struct tah_family_options {
  int32_t order;
  int32_t restriction;
  int32_t threads;  /* workgroup size of its launches: 256 (default), 128 or 64; a tile's units and
                       variables fill that many threads */
  int32_t merge;    /* every unit class of the family in one launch (1) or one launch each (0) */
  int32_t pairs;    /* prolongation over pairs of neighbouring variables where variables are innermost
                       (1; 16-byte loads where the count is even and the array aligned, else pairs of
                       single loads, the last variable of an odd count alone) or one variable per
                       thread (0); 0 for restriction. The plan-wide interp_pairs sets it for even
                       counts only: over an odd count pairs measured slower */
  int32_t lds;      /* restriction streamed along z through LDS (2, any layout), from LDS copies of the
                       source boxes chunk by chunk where the layout allows (1), or the unit kernels (0) */
  int32_t vars;     /* where the layout holds variables innermost: the variables of a tile, 0 (default)
                       for every variable, else at most this many (even, over pairs), the tile's other
                       threads over more units */
} tah_family_options;
/* The families the plan's interpolations fall in and their options: call
 * with out NULL for the count in *n, then with an array of *n entries. */
tah_status tah_plan_get_family_options(const tah_plan* p, int32_t* n, tah_family_options* out);
/* One family's options (named by order and restriction); rebuilds the
 * plan's tiles (it waits for the device first). */
tah_status tah_plan_set_family_options(tah_plan* p, const tah_family_options* o);

/* The autotuner: times every combination of the plan's options on `work`
 * (fills it a few dozen times: a fill is idempotent) and keeps the fastest;
 * then, from there, each interpolation family's options in turn (workgroup
 * size, merged launch, variable pairs, LDS staging), each judged by the
 * time of that family's own launches, and keeps the family choices where
 * the whole fill is no slower with them. With cache_path, the choice is
 * appended to that text file under a signature of the mesh, layout and
 * device, and a later call with the same signature takes it from there
 * without timing (best_ms then -1). best_ms (may be NULL): the chosen
 * options' median fill time. */
tah_status tah_tune(tah_plan* p, double* work, const double* factors, const char* cache_path,
                    double* best_ms);
/* The autotuner of a batched fill (no counterpart in the reference, which
 * fills one field set at a time): times tah_fill_batched over the batch's
 * arrays for every combination of the plan-wide options fork, defer and
 * copy_per_thread, the same on every plan (the other options stay each
 * plan's own), and sets the fastest on every plan. Cached as tah_tune's
 * choice is, under a signature of every plan of the batch; best_ms as
 * tah_tune's. */
tah_status tah_tune_batched(tah_ctx* ctx, int32_t n, tah_plan* const* plans, double* const* works,
                            const double* const* factors, const char* cache_path, double* best_ms);

/* --- The state vector (src/state.jl) -------------------------------- */

/* A state vector holds every own point of every block, no ghosts, laid out
 * (x, y, z, variable, block) with x fastest (statearray,
 * src/state.jl:54-58). The own range of a block is stored points
 * first[d] .. first[d]+extent[d]-1 per dimension: first = G and extent =
 * N + c_d (c_d = 1 along a vertex-like dimension) for d < dim, first = 0 and
 * extent = 1 beyond (ownextents, src/storage.jl:939).
 * statelength (src/state.jl:25): prod(extent) * nvars * nblocks. */
tah_status tah_state_length(const tah_layout* layout, const int32_t* extent, int64_t* out);
/* scatter!(fs, u) (src/state.jl:135-139): state -> the working array's own
 * range; ghosts untouched. Asynchronous on `stream`. */
tah_status tah_scatter(const tah_layout* layout, const int32_t* first, const int32_t* extent,
                       double* work, const double* state, void* stream);
/* gather!(u, fs) (src/state.jl:150-154): the working array's own range ->
 * state. Asynchronous on `stream`. */
tah_status tah_gather(const tah_layout* layout, const int32_t* first, const int32_t* extent,
                      double* state, const double* work, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TAH_H */
