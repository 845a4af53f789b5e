/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY. The CPU restatement of TreeAMR.jl's
 * ghost fill, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker and the CPU baseline. The product path
 * (libtah) never links, loads or calls it.
 *
 * What it restates, on the flattened schedule of include/tah.h:
 *   - run_phase!(::CPU) (src/ghosts.jl:572-593): a phase is one parallel
 *     loop; thread t owns the contiguous block range equalsplit(nblocks, T, t)
 *     (src/threading.jl:50-70) and runs, for every group, the transfers whose
 *     target block it owns, found by bisection of the non-decreasing
 *     targetblocks;
 *   - transfer_kernel! (src/ghosts.jl:372-402) per written value: a unit copy
 *     is 0 + u (copy_load, :127-128); otherwise stencil_sum (:58-118): one
 *     accumulator from +0, m1 innermost, the tap weight formed as
 *     ((w1*w2)*w3) and added as acc += w*u, with the CPU's one-tap skip on a
 *     column holding a single nonzero weight (single_rows,
 *     src/schedule.jl:86-89); then scaled (:369-370): acc*f + 0 when the group
 *     has a factor column;
 *   - exchange_ghosts! (src/ghosts.jl:1099-1160): stages in order (P, 0,
 *     phase 1, prolongation levels coarsest first), and with the outer-face
 *     hook (oracle_fill_boundary) the hook between the stage groups, its
 *     values given by the caller.
 * Built with -ffp-contract=off so that no product and sum fuse.
 *
 * Also: oracle_fill_adjoint, the transpose of the same linear map assembled
 * on the CPU from the schedule (stages reversed; within a stage the targets
 * are read and zeroed, then every source accumulates w*f*adj[target] from
 * the values read, so that a transfer onto itself -- stage P -- scales). It has
 * no counterpart in the reference (which has no autodiff); it is pinned by
 * the dot-product identity <fill(x), y> = <x, fill^T(y)> in tests/.
 *
 * And oracle_scatter / oracle_gather, scatter_kernel! and gather_kernel!
 * (src/state.jl:101-115) as loops over the state vector's linear index n,
 * whose position (i1, i2, i3, var, block) in the column-major shape
 * (extent..., nvars, nblocks) picks the working array's stored point
 * i_d + first_d (first_d = G_d).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/tah.h"

typedef struct {
  int32_t lo, hi; /* 0-based, half-open */
} span_t;

/* src/threading.jl:66-70, 0-based and half-open. */
static span_t equalsplit(int32_t n, int32_t p, int32_t i) {
  int32_t len = n / p, extra = n % p;
  int32_t lo = i * len + (i < extra ? i : extra);
  span_t s = {lo, lo + len + (i < extra ? 1 : 0)};
  return s;
}

/* The transfers of a group whose target block lies in [lo, hi):
 * searchsortedfirst / searchsortedlast on targetblocks. */
static span_t owned_transfers(const tah_group_desc* g, int32_t lo, int32_t hi) {
  int32_t a = 0, b = g->ntransfers;
  while (a < b) {
    int32_t m = (a + b) / 2;
    if (g->targetblocks[m] < lo) a = m + 1; else b = m;
  }
  int32_t first = a;
  a = first;
  b = g->ntransfers;
  while (a < b) {
    int32_t m = (a + b) / 2;
    if (g->targetblocks[m] < hi) a = m + 1; else b = m;
  }
  span_t s = {first, a};
  return s;
}

/* The row of a column's single nonzero weight (0-based), or -1. */
static int single_row(const tah_stencil1d* s, int32_t col) {
  int nz = 0, row = -1;
  for (int m = 0; m < s->order; ++m)
    if (s->weights[(int64_t)s->order * col + m] != 0.0) { ++nz; row = m; }
  return nz == 1 ? row : -1;
}

static void run_transfers(const tah_group_desc* g, const tah_layout* L, double* work, const double* src,
                          const double* factors, int32_t t0, int32_t t1, int skip) {
  const tah_stencil1d* S = g->stencils;
  const int64_t* st = L->stride;
  int unit = S[0].unit && S[1].unit && S[2].unit && S[0].order == 1 && S[1].order == 1 &&
             S[2].order == 1;
  int32_t n0 = S[0].ntarget, n1 = S[1].ntarget, n2 = S[2].ntarget;
  for (int32_t t = t0; t < t1; ++t) {
    double* tb = work + st[4] * g->targetblocks[t];
    const double* sb = src + st[4] * g->sourceblocks[t];
    for (int32_t v = 0; v < L->nvars; ++v) {
      double f = g->factorcol ? factors[v + (int64_t)L->nvars * (g->factorcol - 1)] : 1.0;
      for (int32_t k = 0; k < n2; ++k)
        for (int32_t j = 0; j < n1; ++j)
          for (int32_t i = 0; i < n0; ++i) {
            int32_t b0 = S[0].srcstart[i], b1 = S[1].srcstart[j], b2 = S[2].srcstart[k];
            double acc;
            if (unit) {
              acc = 0.0 + sb[st[0] * b0 + st[1] * b1 + st[2] * b2 + st[3] * v];
            } else {
              int lo0 = 0, hi0 = S[0].order, lo1 = 0, hi1 = S[1].order, lo2 = 0, hi2 = S[2].order;
              if (skip) {
                int r;
                if (S[0].order > 1 && (r = single_row(&S[0], i)) >= 0) { lo0 = r; hi0 = r + 1; }
                if (S[1].order > 1 && (r = single_row(&S[1], j)) >= 0) { lo1 = r; hi1 = r + 1; }
                if (S[2].order > 1 && (r = single_row(&S[2], k)) >= 0) { lo2 = r; hi2 = r + 1; }
              }
              const double* w0 = S[0].weights + (int64_t)S[0].order * i;
              const double* w1 = S[1].weights + (int64_t)S[1].order * j;
              const double* w2 = S[2].weights + (int64_t)S[2].order * k;
              acc = 0.0;
              for (int m2 = lo2; m2 < hi2; ++m2)
                for (int m1 = lo1; m1 < hi1; ++m1)
                  for (int m0 = lo0; m0 < hi0; ++m0) {
                    double w = (w0[m0] * w1[m1]) * w2[m2];
                    acc += w * sb[st[0] * (b0 + m0) + st[1] * (b1 + m1) + st[2] * (b2 + m2) + st[3] * v];
                  }
            }
            if (g->factorcol) acc = acc * f + 0.0;
            tb[st[0] * (S[0].targetfirst + i) + st[1] * (S[1].targetfirst + j) +
               st[2] * (S[2].targetfirst + k) + st[3] * v] = acc;
          }
    }
  }
}

/* One stage of a fill as one parallel loop over T threads. */
static void run_stage(const tah_schedule_desc* d, const tah_layout* L, double* work, const double* factors,
                      int32_t stage, int T, int skip) {
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    int nt = omp_get_num_threads();
    span_t own = equalsplit(d->nblocks, nt, t);
    for (int32_t gi = 0; gi < d->ngroups; ++gi) {
      const tah_group_desc* g = &d->groups[gi];
      if (g->stage != stage) continue;
      span_t tr = owned_transfers(g, own.lo, own.hi);
      if (tr.lo < tr.hi) run_transfers(g, L, work, work, factors, tr.lo, tr.hi, skip);
    }
  }
}

static int fill_threads(const tah_schedule_desc* d, int nthreads) {
  int T = nthreads > 0 ? nthreads : omp_get_max_threads();
  if (T > d->nblocks) T = d->nblocks > 0 ? d->nblocks : 1;
  return T;
}

/* One full fill. nthreads <= 0: all OpenMP threads. skip: the CPU's one-tap
 * skip (1) or the device's full sum (0). */
int oracle_fill(const tah_schedule_desc* d, const tah_layout* L, double* work,
                const double* factors, int nthreads, int skip) {
  if (!d || !L || !work) return 1;
  for (int32_t gi = 0; gi < d->ngroups; ++gi)
    if (d->groups[gi].factorcol && !factors) return 2;
  const int T = fill_threads(d, nthreads);
  for (int32_t stage = 0; stage < d->nstages; ++stage) run_stage(d, L, work, factors, stage, T, skip);
  return 0;
}

/* The hook over the regions of one kind: every region (relative = 0, the
 * values form: boundary_kernel!, src/ghosts.jl:1177-1200), or the regions
 * whose slot is the tag of a stage of stage group k (relative = 1:
 * relative_kernel!, :1240-1260, which sets each ghost to f(x_g) + (u(r) -
 * f(x_r)), r the ghost's point with its clamped components replaced, as
 * relative_point does, :1239). Point p of a region is (i, j, k) of its box
 * with i fastest, numbered from the region's `first`; its values sit at
 * p*nvars + v. */
static void apply_hook(const tah_schedule_desc* d, const tah_layout* L, double* work, int relative,
                       const double* values, const double* fr, int32_t k) {
  const int32_t nv = L->nvars;
  for (int32_t ri = 0; ri < d->nboundaries; ++ri) {
    const tah_boundary_region* R = &d->boundaries[ri];
    if (relative) {
      int in = 0;
      for (int32_t st = d->stagegroup_first[k]; st <= d->stagegroup_last[k]; ++st)
        in = in || d->stage_tags[st] == R->slot;
      if (!in) continue;
    }
    const int32_t n0 = R->hi[0] - R->lo[0], n1 = R->hi[1] - R->lo[1];
    double* b = work + L->stride[4] * R->block;
    for (int32_t z = R->lo[2]; z < R->hi[2]; ++z)
      for (int32_t y = R->lo[1]; y < R->hi[1]; ++y)
        for (int32_t x = R->lo[0]; x < R->hi[0]; ++x) {
          const int64_t p =
              R->first + ((int64_t)(z - R->lo[2]) * n1 + (y - R->lo[1])) * n0 + (x - R->lo[0]);
          const int32_t rx = R->clamps[0] >= 0 ? R->clamps[0] : x;
          const int32_t ry = R->clamps[1] >= 0 ? R->clamps[1] : y;
          const int32_t rz = R->clamps[2] >= 0 ? R->clamps[2] : z;
          for (int32_t v = 0; v < nv; ++v) {
            const int64_t at = p * nv + v;
            double u = values[at];
            if (relative)
              u = u + (b[L->stride[0] * rx + L->stride[1] * ry + L->stride[2] * rz + L->stride[3] * v] - fr[at]);
            b[L->stride[0] * x + L->stride[1] * y + L->stride[2] * z + L->stride[3] * v] = u;
          }
        }
  }
}

/* fill_ghosts!(fs, schedule; boundary) (exchange_ghosts!,
 * src/ghosts.jl:1099-1160): the stage groups in order; after phase 1's
 * group the values hook over every region (kind 0; apply_boundary!,
 * :1134-1136), or after each group the relative hook over the regions whose
 * slot stage it holds (kind 1; phase 1's through apply_boundary!, the
 * others through boundary_after_stage!, :1151-1156). */
int oracle_fill_boundary(const tah_schedule_desc* d, const tah_layout* L, double* work, const double* factors,
                         int kind, const double* values, const double* fr, int nthreads) {
  if (!d || !L || !work) return 1;
  for (int32_t gi = 0; gi < d->ngroups; ++gi)
    if (d->groups[gi].factorcol && !factors) return 2;
  if (kind != 0 && kind != 1) return 3;
  if (d->boundary_points > 0 && (!values || (kind == 1 && !fr))) return 4;
  const int T = fill_threads(d, nthreads);
  for (int32_t k = 0; k < d->nstagegroups; ++k) {
    int phase1 = 0;
    for (int32_t st = d->stagegroup_first[k]; st <= d->stagegroup_last[k]; ++st) {
      run_stage(d, L, work, factors, st, T, 1);
      phase1 = phase1 || d->stage_tags[st] == TAH_STAGE_PHASE1;
    }
    if (kind == 0 && phase1) apply_hook(d, L, work, 0, values, NULL, k);
    if (kind == 1) apply_hook(d, L, work, 1, values, fr, k);
  }
  return 0;
}

/* The regrid transfer (run_stages! in regrid!, src/regrid.jl:1337-1384):
 * stage REGRID reads old_work (the same layout per block as L, the
 * schedule's nsources blocks) and writes new_work; stage REGRID_KEPT copies
 * within new_work. Threads own the new blocks, as the reference threads the
 * transfer by owner of the new block. */
int oracle_regrid(const tah_schedule_desc* d, const tah_layout* L, const double* old_work, double* new_work,
                  int nthreads) {
  if (!d || !L || !old_work || !new_work) return 1;
  int T = nthreads > 0 ? nthreads : omp_get_max_threads();
  if (T > d->nblocks) T = d->nblocks > 0 ? d->nblocks : 1;
  for (int32_t stage = 0; stage < d->nstages; ++stage) {
    const double* src = d->stage_tags[stage] == TAH_STAGE_REGRID ? old_work : new_work;
#pragma omp parallel num_threads(T)
    {
      int t = omp_get_thread_num();
      int nt = omp_get_num_threads();
      span_t own = equalsplit(d->nblocks, nt, t);
      for (int32_t gi = 0; gi < d->ngroups; ++gi) {
        const tah_group_desc* g = &d->groups[gi];
        if (g->stage != stage) continue;
        span_t tr = owned_transfers(g, own.lo, own.hi);
        if (tr.lo < tr.hi) run_transfers(g, L, new_work, src, NULL, tr.lo, tr.hi, 1);
      }
    }
  }
  return 0;
}

/* The transpose of oracle_fill, serial, scatter form. */
int oracle_fill_adjoint(const tah_schedule_desc* d, const tah_layout* L, double* adj,
                        const double* factors) {
  if (!d || !L || !adj) return 1;
  const int64_t* st = L->stride;
  for (int32_t stage = d->nstages - 1; stage >= 0; --stage) {
    /* The stage's target values before it (a transfer of stage P reads its
     * own target), then the targets zeroed: they were overwritten. */
    int64_t total = 0;
    for (int32_t gi = 0; gi < d->ngroups; ++gi) {
      const tah_group_desc* g = &d->groups[gi];
      if (g->stage != stage) continue;
      if (g->factorcol && !factors) return 2;
      const tah_stencil1d* S = g->stencils;
      total += (int64_t)g->ntransfers * L->nvars * S[0].ntarget * S[1].ntarget * S[2].ntarget;
    }
    double* old = (double*)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
    if (!old) return 3;
    int64_t n = 0;
    for (int pass = 0; pass < 2; ++pass) {
      n = 0;
      for (int32_t gi = 0; gi < d->ngroups; ++gi) {
        const tah_group_desc* g = &d->groups[gi];
        if (g->stage != stage) continue;
        const tah_stencil1d* S = g->stencils;
        for (int32_t t = 0; t < g->ntransfers; ++t) {
          double* tb = adj + st[4] * g->targetblocks[t];
          for (int32_t v = 0; v < L->nvars; ++v)
            for (int32_t k = 0; k < S[2].ntarget; ++k)
              for (int32_t j = 0; j < S[1].ntarget; ++j)
                for (int32_t i = 0; i < S[0].ntarget; ++i) {
                  double* x = tb + st[0] * (S[0].targetfirst + i) + st[1] * (S[1].targetfirst + j) +
                              st[2] * (S[2].targetfirst + k) + st[3] * v;
                  if (pass == 0) old[n] = *x; else *x = 0.0;
                  ++n;
                }
        }
      }
    }
    /* Every source accumulates w*f*adj[target], from the values before. */
    n = 0;
    for (int32_t gi = 0; gi < d->ngroups; ++gi) {
      const tah_group_desc* g = &d->groups[gi];
      if (g->stage != stage) continue;
      const tah_stencil1d* S = g->stencils;
      for (int32_t t = 0; t < g->ntransfers; ++t) {
        double* sb = adj + st[4] * g->sourceblocks[t];
        for (int32_t v = 0; v < L->nvars; ++v) {
          double f = g->factorcol ? factors[v + (int64_t)L->nvars * (g->factorcol - 1)] : 1.0;
          for (int32_t k = 0; k < S[2].ntarget; ++k)
            for (int32_t j = 0; j < S[1].ntarget; ++j)
              for (int32_t i = 0; i < S[0].ntarget; ++i) {
                double a = old[n++] * f;
                const double* w0 = S[0].weights + (int64_t)S[0].order * i;
                const double* w1 = S[1].weights + (int64_t)S[1].order * j;
                const double* w2 = S[2].weights + (int64_t)S[2].order * k;
                for (int m2 = 0; m2 < S[2].order; ++m2)
                  for (int m1 = 0; m1 < S[1].order; ++m1)
                    for (int m0 = 0; m0 < S[0].order; ++m0) {
                      double w = (w0[m0] * w1[m1]) * w2[m2];
                      sb[st[0] * (S[0].srcstart[i] + m0) + st[1] * (S[1].srcstart[j] + m1) +
                         st[2] * (S[2].srcstart[k] + m2) + st[3] * v] += w * a;
                    }
              }
        }
      }
    }
    free(old);
  }
  return 0;
}

/* How many times each stored point is written by one fill (the partition
 * oracle write_counts, test/ghost_oracles.jl:228-255): counts has one int32
 * per stored spatial point per block (variables share a box). */
int oracle_write_counts(const tah_schedule_desc* d, const tah_layout* L, int32_t* counts) {
  int64_t per = (int64_t)L->stored[0] * L->stored[1] * L->stored[2];
  for (int32_t gi = 0; gi < d->ngroups; ++gi) {
    const tah_group_desc* g = &d->groups[gi];
    const tah_stencil1d* S = g->stencils;
    for (int32_t t = 0; t < g->ntransfers; ++t)
      for (int32_t k = 0; k < S[2].ntarget; ++k)
        for (int32_t j = 0; j < S[1].ntarget; ++j)
          for (int32_t i = 0; i < S[0].ntarget; ++i)
            counts[per * g->targetblocks[t] +
                   (S[0].targetfirst + i) +
                   (int64_t)L->stored[0] * ((S[1].targetfirst + j) + (int64_t)L->stored[1] * (S[2].targetfirst + k))] += 1;
  }
  return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* The working array's element of state entry n (src/state.jl:101-115). */
static int64_t state_to_work(const tah_layout* L, const int32_t* first, const int32_t* ext, int64_t n) {
  int64_t r = n;
  int64_t i[3];
  for (int d = 0; d < 3; ++d) {
    i[d] = r % ext[d];
    r /= ext[d];
  }
  const int64_t v = r % L->nvars, b = r / L->nvars;
  return L->stride[4] * b + L->stride[3] * v + L->stride[0] * (i[0] + first[0]) +
         L->stride[1] * (i[1] + first[1]) + L->stride[2] * (i[2] + first[2]);
}

static int64_t state_length(const tah_layout* L, const int32_t* ext) {
  return (int64_t)ext[0] * ext[1] * ext[2] * L->nvars * L->nblocks;
}

int oracle_scatter(const tah_layout* L, const int32_t* first, const int32_t* ext, double* work,
                   const double* state) {
  const int64_t n = state_length(L, ext);
  for (int64_t k = 0; k < n; ++k) work[state_to_work(L, first, ext, k)] = state[k];
  return 0;
}

int oracle_gather(const tah_layout* L, const int32_t* first, const int32_t* ext, double* state,
                  const double* work) {
  const int64_t n = state_length(L, ext);
  for (int64_t k = 0; k < n; ++k) state[k] = work[state_to_work(L, first, ext, k)];
  return 0;
}
