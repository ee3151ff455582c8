/*
 * gvdb_oracle.c -- CPU restatement of the GVDB point-to-sparse-grid path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg; never by the product library.
 *
 * Restated from the reference (paths relative to the gvdb-voxels tree; abbreviations
 * in gvdb_oracle.h).  Integer/topology rules follow the reference's FP32 conversions
 * exactly; channel values are FP64.
 *
 * Defined-away reference quirks (each is undefined or racy behaviour in the reference
 * and never reached by the configured workloads):
 *   - lookups start at the root (level = depth); the reference starts at "top_lev", the
 *     smallest level holding exactly one node (V.cpp:3955-3967), which differs only when
 *     a level below the root holds a single node;
 *   - positions outside the root are "missing" (N.cuh:275 checks '>' only and never
 *     checks negatives, so such lookups read outside the child list);
 *   - extra/incremental box nodes outside the root are not emitted.
 *
 * Pinning: every function below is checked against an independent restatement in tests/
 * (set enumeration, closed forms, brute force, hand-computed words, finite differences):
 *   orc_rebuild_topology (full, extra, incremental, state machine, depth errors) and
 *     orc_update_atlas .............................. tests/test_oracle_topology.py
 *   orc_accumulate_topology ............................ tests/test_oracle_pins.py
 *   orc_insert_subcell, orc_gather_density / _levelset (plain), orc_gather_velocity (plain),
 *     orc_update_apron, orc_sample_trilinear, orc_splat_trilinear, orc_splat_backward,
 *     orc_sample_backward, orc_scatter_density, orc_gather_colour, orc_scatter_colour,
 *     orc_quantize_list_pos, orc_gather_levelset_q16 ... tests/test_oracle_fields.py
 *   the accumulating forms of orc_gather_density / _levelset / _velocity,
 *     orc_quantize_list_vel, orc_scatter_reduce_levelset .. tests/test_oracle_pins.py
 *   orc_cpu_activate_space ............................. tests/test_oracle_topology.py
 *   the threaded forms (orc_set_threads > 1) against the serial ones, list for list and
 *     voxel for voxel ................................... tests/test_oracle_threads.py
 * No function is left unpinned.
 */
#include "gvdb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the binning, gather and apron loops (orc_set_threads; 1 = the serial loops).  The
 * threaded forms give the same results as the serial ones: the insert places each thread's
 * contiguous chunk of points after the chunks before it, so every list keeps point order, and
 * every other threaded loop writes disjoint voxels. */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

struct orc_tree {
  int levs;
  int ld[ORC_MAXLEV];
  orc_node* nodes[ORC_MAXLEV];
  int64_t cnt[ORC_MAXLEV], cap[ORC_MAXLEV];
  uint64_t* clist[ORC_MAXLEV];
  int64_t ccnt[ORC_MAXLEV], ccap[ORC_MAXLEV];
  int64_t used0;
  int rebuild;     /* mRebuildTopo */
  int cur_depth;   /* mCurrDepth */
  int depth;
  int32_t rootpos[3];
  float posmin[3], posmax[3];
  int cx, cy, apron;
  /* atlas */
  int64_t nslots;
  int64_t* leaf_slot; int64_t leaf_slot_cap;
  double* chan[32];
  int64_t chan_cap;
  /* subcells */
  int sc, scdim;
  int64_t nsc, sclen;
  int32_t* sc_map; int32_t* sc_cnt; int32_t* sc_off; int32_t* sc_pos; int32_t* sc_nid;
  int64_t* sc_list; float* sc_list_pos;
};

/* ---------------------------------------------------------------- helpers */

static int res_of(const orc_tree* t, int l) { return 1 << t->ld[l]; }

/* getRange (V.h:614-619): product of per-level resolutions up to lev. */
static int64_t range_of(const orc_tree* t, int l) {
  int64_t r = 1;
  for (int k = 0; k <= l && k < t->levs; k++) r <<= t->ld[k];
  return r;
}

static uint64_t elem(int grp, int lev, uint64_t ndx) { return (uint64_t)grp | ((uint64_t)lev << 8) | (ndx << 16); }

static void* grow(void* p, int64_t* cap, int64_t need, size_t esz) {
  if (need <= *cap) return p;
  int64_t c = *cap ? *cap : 4;
  while (c < need) c *= 2;
  void* q = realloc(p, (size_t)c * esz);
  memset((char*)q + (size_t)(*cap) * esz, 0, (size_t)(c - *cap) * esz);
  *cap = c;
  return q;
}

/* PoolAlloc (A.cpp:163-192) for group 0 + SetupNode (V.cpp:2512-2524). */
static int64_t alloc_node(orc_tree* t, int lev, int32_t x, int32_t y, int32_t z) {
  t->nodes[lev] = (orc_node*)grow(t->nodes[lev], &t->cap[lev], t->cnt[lev] + 1, sizeof(orc_node));
  int64_t i = t->cnt[lev]++;
  orc_node* nd = &t->nodes[lev][i];
  memset(nd, 0, sizeof(*nd));
  nd->lev = (uint8_t)lev;
  nd->flags = 1;
  nd->pos[0] = x; nd->pos[1] = y; nd->pos[2] = z;
  nd->value[0] = nd->value[1] = nd->value[2] = -1;
  nd->parent = ORC_UNDEFL;
  nd->childlist = ORC_UNDEFL;
  if (lev == 0) t->used0++;
  if (lev > 0) {
    int64_t n = (int64_t)1 << (3 * t->ld[lev]);
    t->clist[lev] = (uint64_t*)grow(t->clist[lev], &t->ccap[lev], (t->ccnt[lev] + 1) * n, sizeof(uint64_t));
    int64_t ci = t->ccnt[lev]++;
    for (int64_t k = 0; k < n; k++) t->clist[lev][ci * n + k] = ORC_UNDEF64;
    nd->childlist = elem(1, lev, (uint64_t)ci);
  }
  return i;
}

/* Clear (V.cpp:2487-2503): empty pools, keep memory. */
static void clear_tree(orc_tree* t) {
  for (int l = 0; l < t->levs; l++) { t->cnt[l] = 0; t->ccnt[l] = 0; }
  t->used0 = 0;
  t->rebuild = 1;
}

orc_tree* orc_create(int levs, const int* ld, int cx, int cy, int apron) {
  if (levs < 2 || levs > ORC_MAXLEV) return NULL;
  orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
  t->levs = levs;
  for (int l = 0; l < levs; l++) t->ld[l] = ld[l] == 0 ? 1 : ld[l];   /* Configure V.cpp:2385-2388 */
  t->cx = cx; t->cy = cy; t->apron = apron;
  t->rebuild = 1;      /* Initialize V.cpp:2355-2356 */
  t->cur_depth = -1;
  t->depth = -1;
  return t;
}

void orc_destroy(orc_tree* t) {
  if (!t) return;
  for (int l = 0; l < ORC_MAXLEV; l++) { free(t->nodes[l]); free(t->clist[l]); }
  for (int c = 0; c < 32; c++) free(t->chan[c]);
  free(t->leaf_slot);
  free(t->sc_map); free(t->sc_cnt); free(t->sc_off); free(t->sc_pos); free(t->sc_nid);
  free(t->sc_list); free(t->sc_list_pos);
  free(t);
}

void orc_request_full_rebuild(orc_tree* t, int f) { t->rebuild = f ? 1 : 0; }

/* Tree descent from the root to level stop_lev (getPosParent N.cuh:303-348,
 * getPosLeafParent N.cuh:255-288, getParent N.cuh:350-397).  Returns the pool index
 * of the linked node at stop_lev containing integer position p, or -1. */
static int64_t descend(const orc_tree* t, const int32_t* p, int stop_lev) {
  int d = t->depth;
  if (d < 0 || t->cnt[d] < 1) return -1;
  int64_t R = range_of(t, d);
  for (int a = 0; a < 3; a++)
    if ((int64_t)p[a] < t->rootpos[a] || (int64_t)p[a] >= t->rootpos[a] + R) return -1;
  if (stop_lev == d) return 0;
  int64_t cur = 0;
  for (int l = d; l > stop_lev; l--) {
    const orc_node* nd = &t->nodes[l][cur];
    int res = res_of(t, l);
    int64_t cr = range_of(t, l) / res;   /* child range */
    int64_t b[3];
    for (int a = 0; a < 3; a++) b[a] = ((int64_t)p[a] - nd->pos[a]) / cr;
    int64_t bit = (b[2] * res + b[1]) * res + b[0];
    if (nd->childlist == ORC_UNDEFL) return -1;
    uint64_t ci = nd->childlist >> 16;
    uint64_t e = t->clist[l][ci * ((uint64_t)res * res * res) + (uint64_t)bit];
    if (e == ORC_UNDEF64) return -1;
    cur = (int64_t)(e >> 16);
  }
  return cur;
}

int64_t orc_find_leaf(const orc_tree* t, const int32_t* p) { return descend(t, p, 0); }

/* gvdbLinkBricks (N.cuh:470-505) for every flagged node of one level. */
static void link_level(orc_tree* t, int lev) {
  int res = res_of(t, lev + 1);
  int64_t cr = range_of(t, lev + 1) / res;
  for (int64_t i = 0; i < t->cnt[lev]; i++) {
    orc_node* nd = &t->nodes[lev][i];
    if (!nd->flags) continue;
    int64_t pi = descend(t, nd->pos, lev + 1);
    if (pi < 0) continue;
    orc_node* pn = &t->nodes[lev + 1][pi];
    int64_t b[3];
    for (int a = 0; a < 3; a++) b[a] = ((int64_t)nd->pos[a] - pn->pos[a]) / cr;
    int64_t bit = (b[2] * res + b[1]) * res + b[0];
    nd->parent = elem(0, lev + 1, (uint64_t)pi);
    uint64_t ci = pn->childlist >> 16;
    t->clist[lev + 1][ci * ((uint64_t)res * res * res) + (uint64_t)bit] = elem(0, lev, (uint64_t)i);
  }
}

/* GetBoundingBox (V.cpp:5404-5467) over all points, invalid ones included. */
static void bounding_box(orc_tree* t, const float* pos, int64_t n, int64_t st, const float* tr) {
  float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      float q = pos[i * st + a] + tr[a];
      if (i == 0 || q < mn[a]) mn[a] = q;
      if (i == 0 || q > mx[a]) mx[a] = q;
    }
  }
  for (int a = 0; a < 3; a++) {
    t->posmin[a] = mn[a] - 1.0f;
    t->posmax[a] = mx[a] + 2.0f;
    if (0 > t->posmin[a]) t->posmin[a] = 0;
    if (0 > t->posmax[a]) t->posmax[a] = 0;
  }
}

/* DetermineDepth (V.cpp:821-837) with host GetCoveringNode (V.cpp:2915-2931). */
static int determine_depth(orc_tree* t, int32_t* root) {
  for (int lev = 0; lev < t->levs; lev++) {
    int64_t R = range_of(t, lev);
    int same = 1;
    int32_t c0[3];
    for (int a = 0; a < 3; a++) {
      int64_t pm = (int32_t)t->posmin[a], px = (int32_t)t->posmax[a];
      int64_t cm = pm / R * R, cx = px / R * R;
      if (pm < cm) cm -= R;
      if (px < cx) cx -= R;
      c0[a] = (int32_t)cm;
      if (cm != cx) same = 0;
    }
    if (same) { root[0] = c0[0]; root[1] = c0[1]; root[2] = c0[2]; return lev; }
  }
  return -1;
}

/* GetCoveringNode(float3,int) (P.cuh:88-100): ceil-then-adjust, FP32. */
static int32_t cover_f(float p, int32_t R) {
  float c = ceilf(p / (float)R) * (float)R;
  int32_t np = (int32_t)c;
  if (p < (float)np) np -= R;
  return np;
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

static uint64_t make_key(int lev, int64_t kx, int64_t ky, int64_t kz) {
  return ((uint64_t)lev << 48) | ((uint64_t)(uint16_t)kx << 32) | ((uint64_t)(uint16_t)ky << 16) | (uint64_t)(uint16_t)kz;
}

/* Sort, unique (FindUniqueBrick V.cpp:843-873), then allocate per level in key order
 * (V.cpp:1308-1331 / 975-997).  Keys with level 255 are dropped. */
static void alloc_sorted_keys(orc_tree* t, uint64_t* keys, int64_t m, int d) {
  qsort(keys, (size_t)m, sizeof(uint64_t), cmp_u64);
  for (int lev = 0; lev < d; lev++) {
    int64_t R = range_of(t, lev);
    for (int64_t i = 0; i < m; i++) {
      if (i > 0 && keys[i] == keys[i - 1]) continue;
      int kl = (int)((keys[i] >> 48) & 0xFF);
      if (kl != lev) continue;
      int64_t kx = (keys[i] >> 32) & 0xFFFF, ky = (keys[i] >> 16) & 0xFFFF, kz = keys[i] & 0xFFFF;
      alloc_node(t, lev, (int32_t)(kx * R), (int32_t)(ky * R), (int32_t)(kz * R));
    }
  }
}

/* ActivateBricksGPU (V.cpp:1229-1413) with gvdbCalcBrickId (P.cuh:352-382). */
static void activate_bricks(orc_tree* t, const float* pos, int64_t n, int64_t st, const float* tr, int d) {
  uint64_t* keys = (uint64_t*)malloc((size_t)(n * d + 1) * sizeof(uint64_t));
  for (int64_t i = 0; i < n; i++) {
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = pos[i * st + a] + tr[a];
    int inval = (q[0] < 0 || q[1] < 0 || q[2] < 0);
    for (int lev = 0; lev < d; lev++) {
      if (inval) { keys[i * d + lev] = (uint64_t)255 << 48; continue; }
      int32_t R = (int32_t)range_of(t, lev);
      keys[i * d + lev] = make_key(lev, cover_f(q[0], R) / R, cover_f(q[1], R) / R, cover_f(q[2], R) / R);
    }
  }
  /* root (V.cpp:1288-1300) */
  alloc_node(t, d, t->rootpos[0], t->rootpos[1], t->rootpos[2]);
  alloc_sorted_keys(t, keys, n * d, d);
  free(keys);
  for (int lev = d - 1; lev >= 0; lev--) link_level(t, lev);
}

/* Box enumeration shared by gvdbCalcExtraBrickId (P.cuh:467-512, hi = tq + 2r) and
 * gvdbCalcIncreExtraBrickId (P.cuh:420-465, hi = tq + r).  Emits keys of missing nodes;
 * marks found leaves when markers != NULL. */
static int64_t box_keys(orc_tree* t, const float* pos, int64_t n, int64_t st, const float* tr, float r,
                        float hi_mul, int d, uint64_t** keys_io, int64_t* kcap, int32_t* markers) {
  int64_t m = 0;
  int64_t Rd = range_of(t, d);
  for (int64_t i = 0; i < n; i++) {
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = pos[i * st + a] + tr[a];
    if (q[0] < 0 || q[1] < 0 || q[2] < 0) continue;
    int32_t bmin[3], bmax[3];
    for (int a = 0; a < 3; a++) {
      float tq = (float)(int32_t)q[a];
      bmin[a] = (int32_t)(tq - r);
      bmax[a] = (int32_t)(tq + hi_mul * r);
    }
    for (int lev = 0; lev < d; lev++) {
      int32_t R = (int32_t)range_of(t, lev);
      int32_t lo[3], hi[3];
      for (int a = 0; a < 3; a++) { lo[a] = bmin[a] / R * R; hi[a] = bmax[a] / R * R; }
      int32_t p[3];
      for (p[0] = lo[0]; p[0] <= hi[0]; p[0] += R)
        for (p[1] = lo[1]; p[1] <= hi[1]; p[1] += R)
          for (p[2] = lo[2]; p[2] <= hi[2]; p[2] += R) {
            int out = 0;
            for (int a = 0; a < 3; a++)
              if ((int64_t)p[a] < t->rootpos[a] || (int64_t)p[a] >= t->rootpos[a] + Rd) out = 1;
            if (out) continue;
            int64_t f = descend(t, p, lev);
            if (f < 0) {
              *keys_io = (uint64_t*)grow(*keys_io, kcap, m + 1, sizeof(uint64_t));
              (*keys_io)[m++] = make_key(lev, p[0] / R, p[1] / R, p[2] / R);
            } else if (lev == 0 && markers) {
              markers[f] = 1;
            }
          }
    }
  }
  return m;
}

/* ActivateExtraBricksGPU (V.cpp:931-1016). */
static void activate_extra(orc_tree* t, const float* pos, int64_t n, int64_t st, const float* tr, float r, int d) {
  uint64_t* keys = NULL; int64_t kcap = 0;
  int64_t m = box_keys(t, pos, n, st, tr, r, 2.0f, d, &keys, &kcap, NULL);
  if (m > 0) alloc_sorted_keys(t, keys, m, d);
  free(keys);
  for (int lev = d - 1; lev >= 0; lev--) link_level(t, lev);
}

/* ActivateIncreBricksGPU (V.cpp:1020-1156). */
static void activate_incre(orc_tree* t, const float* pos, int64_t n, int64_t st, const float* tr, float r, int d,
                           int accum) {
  int64_t total = t->cnt[0];
  if (total == 0) return;
  int32_t* markers = (int32_t*)calloc((size_t)total, sizeof(int32_t));
  if (accum)
    for (int64_t i = 0; i < total; i++) markers[i] = t->nodes[0][i].flags;
  uint64_t* keys = NULL; int64_t kcap = 0;
  int64_t m = box_keys(t, pos, n, st, tr, r, 1.0f, d, &keys, &kcap, markers);
  if (m == 0) { free(keys); free(markers); return; }   /* V.cpp:1066: no new node -> unchanged */
  for (int64_t i = 0; i < total; i++) t->nodes[0][i].flags = (uint8_t)(markers[i] != 0);
  free(markers);
  alloc_sorted_keys(t, keys, m, d);
  free(keys);
  for (int lev = d - 1; lev >= 0; lev--) link_level(t, lev);
  /* gvdbDelinkLeafBricks (N.cuh:401-429) */
  {
    int res = res_of(t, 1);
    int64_t cr = range_of(t, 1) / res;
    for (int64_t i = 0; i < t->cnt[0]; i++) {
      orc_node* nd = &t->nodes[0][i];
      if (nd->flags) continue;
      if (nd->parent == ORC_UNDEF64) continue;
      int64_t pi = descend(t, nd->pos, 1);
      nd->parent = ORC_UNDEF64;
      if (pi < 0) continue;
      orc_node* pn = &t->nodes[1][pi];
      int64_t b[3];
      for (int a = 0; a < 3; a++) b[a] = ((int64_t)nd->pos[a] - pn->pos[a]) / cr;
      uint64_t ci = pn->childlist >> 16;
      t->clist[1][ci * ((uint64_t)res * res * res) + (uint64_t)((b[2] * res + b[1]) * res + b[0])] = ORC_UNDEF64;
    }
  }
  /* gvdbDelinkBricks (N.cuh:431-467), levels 1..d-1 bottom-up */
  for (int lev = 1; lev < d; lev++) {
    int resl = res_of(t, lev);
    int64_t nch = (int64_t)resl * resl * resl;
    int resp = res_of(t, lev + 1);
    int64_t cr = range_of(t, lev + 1) / resp;
    for (int64_t i = 0; i < t->cnt[lev]; i++) {
      orc_node* nd = &t->nodes[lev][i];
      if (nd->parent == ORC_UNDEF64) { nd->flags = 0; continue; }
      uint64_t ci = nd->childlist >> 16;
      int any = 0;
      for (int64_t k = 0; k < nch; k++)
        if (t->clist[lev][ci * nch + k] != ORC_UNDEF64) { any = 1; break; }
      if (any) { nd->flags = 1; continue; }
      nd->flags = 0;
      int64_t pi = descend(t, nd->pos, lev + 1);
      nd->parent = ORC_UNDEF64;
      if (pi < 0) continue;
      orc_node* pn = &t->nodes[lev + 1][pi];
      int64_t b[3];
      for (int a = 0; a < 3; a++) b[a] = ((int64_t)nd->pos[a] - pn->pos[a]) / cr;
      uint64_t pci = pn->childlist >> 16;
      t->clist[lev + 1][pci * ((uint64_t)resp * resp * resp) + (uint64_t)((b[2] * resp + b[1]) * resp + b[0])] =
          ORC_UNDEF64;
    }
  }
  /* used count (V.cpp:1144-1147) */
  int64_t used = t->cnt[0];
  for (int64_t i = 0; i < t->cnt[0]; i++)
    if (!t->nodes[0][i].flags) used--;
  t->used0 = used;
}

static int check_depth(orc_tree* t, int depth) {
  if (depth == 0) return ORC_ERR_DEPTH0;
  if (depth < 0 || depth > 5 || depth >= t->levs) return ORC_ERR_DEPTH_MAX;
  return ORC_OK;
}

/* RebuildTopology (V.cpp:1181-1218). */
int orc_rebuild_topology(orc_tree* t, const float* pos, int64_t n, int64_t st, float r, const float* tr) {
  if (!t || !pos || n <= 0 || st < 3) return ORC_ERR_ARG;
  bounding_box(t, pos, n, st, tr);
  int32_t root[3];
  int depth = determine_depth(t, root);
  int s = check_depth(t, depth);
  if (s) return s;
  if (t->rebuild) {
    clear_tree(t);
  } else if (depth != t->cur_depth) {
    clear_tree(t);
    t->rebuild = 1;
    t->cur_depth = depth;
  }
  if (t->rebuild) {
    t->depth = depth;
    memcpy(t->rootpos, root, sizeof(root));
    activate_bricks(t, pos, n, st, tr, depth);
    if (r >= 1) activate_extra(t, pos, n, st, tr, r, depth);
    t->rebuild = 0;
  } else {
    activate_incre(t, pos, n, st, tr, r, depth, 0);
  }
  return ORC_OK;
}

/* AccumulateTopology (V.cpp:1158-1178). */
int orc_accumulate_topology(orc_tree* t, const float* pos, int64_t n, int64_t st, float r, const float* tr,
                            int min_depth) {
  if (!t || !pos || n <= 0 || st < 3) return ORC_ERR_ARG;
  bounding_box(t, pos, n, st, tr);
  int32_t root[3];
  int depth = determine_depth(t, root);
  if (depth >= 0 && depth < min_depth) depth = min_depth;
  int s = check_depth(t, depth);
  if (s) return s;
  if (t->rebuild) return orc_rebuild_topology(t, pos, n, st, r, tr);
  /* The reference hands the new points' depth and root to the incremental pass over a tree of
   * the old depth (V.cpp:1177), which is undefined when they differ; the pass keeps the tree's
   * depth and root (points outside the root add nothing), as the library does. */
  activate_incre(t, pos, n, st, tr, r, t->depth, 1);
  return ORC_OK;
}

/* UpdateAtlas (V.cpp:2630-2704) + AtlasAlloc/getAtlasPos (A.cpp:692-715): flagged leaves
 * get consecutive slots in pool order; mValue = slot corner in the reference atlas. */
int orc_update_atlas(orc_tree* t) {
  int64_t total = t->cnt[0];
  t->leaf_slot = (int64_t*)grow(t->leaf_slot, &t->leaf_slot_cap, total + 1, sizeof(int64_t));
  int br = res_of(t, 0) + 2 * t->apron;
  int64_t s = 0;
  for (int64_t i = 0; i < total; i++) {
    orc_node* nd = &t->nodes[0][i];
    if (!nd->flags) { t->leaf_slot[i] = -1; continue; }
    int64_t a2 = (int64_t)t->cx * t->cy;
    int64_t z = s / a2, rem = s - z * a2;
    int64_t y = rem / t->cx, x = rem - y * t->cx;
    nd->value[0] = (int32_t)(x * br + t->apron);
    nd->value[1] = (int32_t)(y * br + t->apron);
    nd->value[2] = (int32_t)(z * br + t->apron);
    t->leaf_slot[i] = s++;
  }
  t->nslots = s;
  int64_t per = (int64_t)br * br * br;
  if (s * per > t->chan_cap) {   /* channels are re-sized together */
    int64_t oldcap = t->chan_cap;
    for (int c = 0; c < 32; c++) {
      if (!t->chan[c]) continue;
      t->chan[c] = (double*)realloc(t->chan[c], (size_t)(s * per) * sizeof(double));
      memset(t->chan[c] + oldcap, 0, (size_t)(s * per - oldcap) * sizeof(double));
    }
    t->chan_cap = s * per;
  }
  return ORC_OK;
}

int64_t orc_node_count(const orc_tree* t, int lev) { return (lev >= 0 && lev < t->levs) ? t->cnt[lev] : 0; }
int64_t orc_used0(const orc_tree* t) { return t->used0; }
int64_t orc_clist_count(const orc_tree* t, int lev) { return (lev > 0 && lev < t->levs) ? t->ccnt[lev] : 0; }
const orc_node* orc_nodes(const orc_tree* t, int lev) { return t->nodes[lev]; }
const uint64_t* orc_clists(const orc_tree* t, int lev) { return t->clist[lev]; }
int orc_depth(const orc_tree* t) { return t->depth; }
void orc_rootpos(const orc_tree* t, int32_t* o) { o[0] = t->rootpos[0]; o[1] = t->rootpos[1]; o[2] = t->rootpos[2]; }
void orc_posminmax(const orc_tree* t, float* o) {
  for (int a = 0; a < 3; a++) { o[a] = t->posmin[a]; o[3 + a] = t->posmax[a]; }
}
int64_t orc_num_slots(const orc_tree* t) { return t->nslots; }
const int64_t* orc_leaf_slot(const orc_tree* t) { return t->leaf_slot; }

/* ----------------------------------------------------------------- subcells */

/* Subcell enumeration of gvdbCountSubcell / gvdbInsertSubcell (P.cuh:275-315, 218-273):
 * corners from trunc(q-r)/sc*sc to trunc(q+1+r)/sc*sc.  Calls fn for each hit. */
typedef void (*sc_fn)(orc_tree* t, int64_t idx, int64_t pt, const float* q, void* u);

static void enum_subcells(orc_tree* t, int sc, const float* pos, int64_t i0, int64_t i1, int64_t st, float r,
                          const float* tr, sc_fn fn, void* u) {
  int D = t->scdim;
  for (int64_t i = i0; i < i1; i++) {
    const float* p = pos + i * st;
    if (p[2] == ORC_NOHIT) continue;
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = p[a] + tr[a];
    if (q[0] < 0 || q[1] < 0 || q[2] < 0) continue;
    int32_t lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      float lq = q[a] - r;
      float hq = q[a] + 1.0f;
      hq = hq + r;
      lo[a] = ((int32_t)lq / sc) * sc;
      hi[a] = ((int32_t)hq / sc) * sc;
    }
    int32_t c[3];
    for (c[0] = lo[0]; c[0] <= hi[0]; c[0] += sc)
      for (c[1] = lo[1]; c[1] <= hi[1]; c[1] += sc)
        for (c[2] = lo[2]; c[2] <= hi[2]; c[2] += sc) {
          int64_t leaf = descend(t, c, 0);
          if (leaf < 0) continue;
          const orc_node* nd = &t->nodes[0][leaf];
          int32_t l[3];
          for (int a = 0; a < 3; a++) l[a] = (c[a] - nd->pos[a]) / sc;
          int64_t idx = (int64_t)t->sc_map[leaf] * D * D * D + ((int64_t)l[2] * D + l[1]) * D + l[0];
          fn(t, idx, i, q, u);
        }
  }
}

static void count_fn(orc_tree* t, int64_t idx, int64_t pt, const float* q, void* u) {
  (void)pt; (void)q;
  ((int32_t*)u)[idx]++;
}

typedef struct { int32_t* cursor; } ins_ctx;
static void insert_fn(orc_tree* t, int64_t idx, int64_t pt, const float* q, void* u) {
  ins_ctx* c = (ins_ctx*)u;
  int64_t k = (int64_t)t->sc_off[idx] + c->cursor[idx]++;
  t->sc_list[k] = pt;
  t->sc_list_pos[3 * k + 0] = q[0];
  t->sc_list_pos[3 * k + 1] = q[1];
  t->sc_list_pos[3 * k + 2] = q[2];
}

int orc_insert_subcell(orc_tree* t, int sc, const float* pos, int64_t n, int64_t st, float r, const float* tr,
                       int64_t* len_out) {
  if (!t || sc <= 0 || st < 3) return ORC_ERR_ARG;
  int res0 = res_of(t, 0);
  if (res0 % sc) return ORC_ERR_ARG;
  if (t->used0 == 0) { *len_out = 0; return ORC_OK; }
  t->sc = sc;
  t->scdim = res0 / sc;
  int D = t->scdim;
  int64_t total = t->cnt[0];
  /* gvdbSetFlagSubcell (P.cuh:109-118) + exclusive PrefixSum (V.cpp:4999-5004) */
  free(t->sc_map);
  t->sc_map = (int32_t*)malloc((size_t)(total + 1) * sizeof(int32_t));
  int32_t acc = 0;
  for (int64_t i = 0; i < total; i++) {
    t->sc_map[i] = acc;
    acc += (t->nodes[0][i].parent == ORC_UNDEF64) ? 0 : 1;
  }
  int64_t nsc = (int64_t)D * D * D * t->used0;
  t->nsc = nsc;
  free(t->sc_cnt); free(t->sc_off); free(t->sc_pos); free(t->sc_nid);
  t->sc_cnt = (int32_t*)calloc((size_t)nsc + 1, sizeof(int32_t));
  t->sc_off = (int32_t*)calloc((size_t)nsc + 1, sizeof(int32_t));
  t->sc_pos = (int32_t*)calloc((size_t)nsc * 3 + 3, sizeof(int32_t));
  t->sc_nid = (int32_t*)calloc((size_t)nsc + 1, sizeof(int32_t));
  if (g_threads <= 1) {
    enum_subcells(t, sc, pos, 0, n, st, r, tr, count_fn, t->sc_cnt);
    int64_t o = 0;
    for (int64_t k = 0; k < nsc; k++) { t->sc_off[k] = (int32_t)o; o += t->sc_cnt[k]; }
    t->sclen = o;
    *len_out = o;
    free(t->sc_list); free(t->sc_list_pos);
    t->sc_list = (int64_t*)malloc((size_t)(o + 1) * sizeof(int64_t));
    t->sc_list_pos = (float*)malloc((size_t)(o + 1) * 3 * sizeof(float));
    ins_ctx c;
    c.cursor = (int32_t*)calloc((size_t)nsc + 1, sizeof(int32_t));
    enum_subcells(t, sc, pos, 0, n, st, r, tr, insert_fn, &c);
    free(c.cursor);
  } else {
    /* thread k counts, then inserts, points [n k / T, n (k + 1) / T); its entries of subcell s
     * start after those of threads 0..k-1, so the lists equal the serial ones */
    const int T = g_threads;
    const int64_t row = nsc + 1;
    int32_t* pc = (int32_t*)calloc((size_t)T * (size_t)row, sizeof(int32_t));
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
      const int id = omp_get_thread_num(), nt = omp_get_num_threads();
#else
      const int id = 0, nt = 1;
#endif
      int32_t* mine = pc + (int64_t)id * row;
      const int64_t i0 = n * id / nt, i1 = n * (id + 1) / nt;
      enum_subcells(t, sc, pos, i0, i1, st, r, tr, count_fn, mine);
#pragma omp barrier
#pragma omp for schedule(static)
      for (int64_t k = 0; k < nsc; k++) {
        int32_t sum = 0;
        for (int th = 0; th < nt; th++) {
          const int32_t v = pc[(int64_t)th * row + k];
          pc[(int64_t)th * row + k] = sum;
          sum += v;
        }
        t->sc_cnt[k] = sum;
      }
#pragma omp single
      {
        int64_t o = 0;
        for (int64_t k = 0; k < nsc; k++) { t->sc_off[k] = (int32_t)o; o += t->sc_cnt[k]; }
        t->sclen = o;
        *len_out = o;
        free(t->sc_list); free(t->sc_list_pos);
        t->sc_list = (int64_t*)malloc((size_t)(o + 1) * sizeof(int64_t));
        t->sc_list_pos = (float*)malloc((size_t)(o + 1) * 3 * sizeof(float));
      }
      ins_ctx c;
      c.cursor = mine;
      enum_subcells(t, sc, pos, i0, i1, st, r, tr, insert_fn, &c);
    }
    free(pc);
  }
  /* gvdbCalcSubcellPos (P.cuh:677-694) */
  int64_t per = (int64_t)D * D * D;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < total; i++) {
    const orc_node* nd = &t->nodes[0][i];
    if (nd->parent == ORC_UNDEF64) continue;
    for (int64_t s = 0; s < per; s++) {
      int64_t k = (int64_t)t->sc_map[i] * per + s;
      if (k >= nsc) continue;
      t->sc_pos[3 * k + 0] = nd->pos[0] + (int32_t)(s % D) * sc;
      t->sc_pos[3 * k + 1] = nd->pos[1] + (int32_t)((s / D) % D) * sc;
      t->sc_pos[3 * k + 2] = nd->pos[2] + (int32_t)((s / D / D) % D) * sc;
      t->sc_nid[k] = (int32_t)i;
    }
  }
  return ORC_OK;
}

int64_t orc_sc_num(const orc_tree* t) { return t->nsc; }
const int32_t* orc_sc_map(const orc_tree* t) { return t->sc_map; }
const int32_t* orc_sc_cnt(const orc_tree* t) { return t->sc_cnt; }
const int32_t* orc_sc_off(const orc_tree* t) { return t->sc_off; }
const int32_t* orc_sc_pos(const orc_tree* t) { return t->sc_pos; }
const int32_t* orc_sc_nid(const orc_tree* t) { return t->sc_nid; }
const int64_t* orc_sc_list(const orc_tree* t) { return t->sc_list; }
const float* orc_sc_list_pos(const orc_tree* t) { return t->sc_list_pos; }

/* ----------------------------------------------------------------- channels */

int orc_channel_alloc(orc_tree* t, int chan) {
  if (chan < 0 || chan >= 32) return ORC_ERR_ARG;
  if (!t->chan[chan]) {
    int64_t n = t->chan_cap > 0 ? t->chan_cap : 1;
    if (t->chan_cap == 0) t->chan_cap = 1;
    t->chan[chan] = (double*)calloc((size_t)n, sizeof(double));
  }
  return ORC_OK;
}
double* orc_channel(orc_tree* t, int chan) { return t->chan[chan]; }
int64_t orc_channel_size(const orc_tree* t) {
  int br = res_of(t, 0) + 2 * t->apron;
  return t->nslots * br * br * br;
}
int orc_fill_channel(orc_tree* t, int chan, double v) {
  if (!t->chan[chan]) return ORC_ERR_ARG;
  int64_t n = orc_channel_size(t);
  for (int64_t i = 0; i < n; i++) t->chan[chan][i] = v;
  return ORC_OK;
}

/* Index into the brick-linear atlas: slot, padded local voxel. */
static int64_t atlas_index(const orc_tree* t, int64_t slot, int lx, int ly, int lz) {
  int br = res_of(t, 0) + 2 * t->apron;
  return ((slot * br + lz) * br + ly) * br + lx;
}

/* Per subcell, per voxel: reduce over the subcell list (P.cuh:696-773). mode 0 = density
 * (max, init 0), 1 = level set (min, init 3.0). */
static int gather_common(orc_tree* t, float r, int chan, int accum, int mode) {
  if (!t->chan[chan] || !t->sc_cnt) return ORC_ERR_STATE;
  int sc = t->sc, a = t->apron;
  double* ch = t->chan[chan];
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) if (g_threads > 1)
  for (int64_t k = 0; k < t->nsc; k++) {
    int64_t leaf = t->sc_nid[k];
    const orc_node* nd = &t->nodes[0][leaf];
    int64_t slot = t->leaf_slot[leaf];
    if (slot < 0) continue;
    for (int tz = 0; tz < sc; tz++)
      for (int ty = 0; ty < sc; ty++)
        for (int tx = 0; tx < sc; tx++) {
          int32_t v[3] = {t->sc_pos[3 * k] + tx, t->sc_pos[3 * k + 1] + ty, t->sc_pos[3 * k + 2] + tz};
          int64_t ai = atlas_index(t, slot, v[0] - nd->pos[0] + a, v[1] - nd->pos[1] + a, v[2] - nd->pos[2] + a);
          double val = accum ? ch[ai] : (mode == 0 ? 0.0 : 3.0);
          for (int64_t j = t->sc_off[k]; j < (int64_t)t->sc_off[k] + t->sc_cnt[k]; j++) {
            double dx = (double)t->sc_list_pos[3 * j] - v[0];
            double dy = (double)t->sc_list_pos[3 * j + 1] - v[1];
            double dz = (double)t->sc_list_pos[3 * j + 2] - v[2];
            double c = sqrt(dx * dx + dy * dy + dz * dz);
            if (mode == 0) { double w = (double)r - c; if (w > val) val = w; }
            else { double w = c - (double)r; if (w < val) val = w; }
          }
          ch[ai] = val;
        }
  }
  return ORC_OK;
}

int orc_gather_density(orc_tree* t, float r, int chan, int accum) { return gather_common(t, r, chan, accum, 0); }

/* FP16 subcell lists (gvdbInsertSubcell_fp16, P.cuh:200-207): ushort((v - lo) / range * 65535)
 * in single precision, each operation rounded in that order, saturated to [0, 65535]. */
static uint16_t quant16(float v, float lo, float range) {
  volatile float d = v - lo;
  volatile float q = d / range;
  float f = q * 65535.0f;
  if (!(f >= 0.0f)) return 0;
  if (f >= 65535.0f) return 65535;
  return (uint16_t)f;
}

/* The listed positions quantized to [qmin, qmin + qrange]: out[3 * len] */
int orc_quantize_list_pos(const orc_tree* t, const float* qmin, const float* qrange, uint16_t* out) {
  if (!t->sc_cnt) return ORC_ERR_STATE;
  int64_t len = t->nsc > 0 ? (int64_t)t->sc_off[t->nsc - 1] + t->sc_cnt[t->nsc - 1] : 0;
  for (int64_t j = 0; j < len; j++)
    for (int a = 0; a < 3; a++) out[3 * j + a] = quant16(t->sc_list_pos[3 * j + a], qmin[a], qrange[a]);
  return ORC_OK;
}

/* The listed points' velocities (vel: float3 rows of `st` floats) quantized the same way */
int orc_quantize_list_vel(const orc_tree* t, const float* vel, int64_t st, const float* vmin, const float* vrange,
                          uint16_t* out) {
  if (!t->sc_cnt) return ORC_ERR_STATE;
  int64_t len = t->nsc > 0 ? (int64_t)t->sc_off[t->nsc - 1] + t->sc_cnt[t->nsc - 1] : 0;
  for (int64_t j = 0; j < len; j++)
    for (int a = 0; a < 3; a++) out[3 * j + a] = quant16(vel[t->sc_list[j] * st + a], vmin[a], vrange[a]);
  return ORC_OK;
}

/* gvdbGatherLevelSet_fp16 (P.cuh:775-811): each listed position decoded as
 * u / 65535 * range + lo (single precision, unfused), then the level set min(3, |p - v| - r)
 * without accumulation. q16: the quantized list (3 per entry). */
int orc_gather_levelset_q16(orc_tree* t, float r, int chan, const uint16_t* q16, const float* qmin,
                            const float* qrange) {
  if (!t->chan[chan] || !t->sc_cnt) return ORC_ERR_STATE;
  int64_t len = t->nsc > 0 ? (int64_t)t->sc_off[t->nsc - 1] + t->sc_cnt[t->nsc - 1] : 0;
  float* dec = (float*)malloc((size_t)(len + 1) * 3 * sizeof(float));
  if (!dec) return ORC_ERR_NOMEM;
  for (int64_t j = 0; j < 3 * len; j++) {
    int a = (int)(j % 3);
    volatile float f = (float)q16[j] / 65535.0f;
    volatile float g = f * qrange[a];
    dec[j] = g + qmin[a];
  }
  float* keep = t->sc_list_pos;
  t->sc_list_pos = dec;
  int s = gather_common(t, r, chan, 0, 1);
  t->sc_list_pos = keep;
  free(dec);
  return s;
}
int orc_gather_levelset(orc_tree* t, float r, int chan, int accum) { return gather_common(t, r, chan, accum, 1); }

/* Velocity channels: the reference colour rule (P.cuh:717-731) restated for float3:
 * (accum ? old : 0) + sum of the subcell's binned velocities, divided by (cnt + accum);
 * an empty subcell without accumulation gives 0. */
int orc_gather_velocity(orc_tree* t, const float* vel, int64_t st, int cx, int cy, int cz, int accum) {
  int chans[3] = {cx, cy, cz};
  for (int c = 0; c < 3; c++) if (!t->chan[chans[c]]) return ORC_ERR_STATE;
  if (!t->sc_cnt) return ORC_ERR_STATE;
  int sc = t->sc, a = t->apron;
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) if (g_threads > 1)
  for (int64_t k = 0; k < t->nsc; k++) {
    int64_t leaf = t->sc_nid[k];
    const orc_node* nd = &t->nodes[0][leaf];
    int64_t slot = t->leaf_slot[leaf];
    if (slot < 0) continue;
    double sum[3] = {0, 0, 0};
    for (int64_t j = t->sc_off[k]; j < (int64_t)t->sc_off[k] + t->sc_cnt[k]; j++)
      for (int c = 0; c < 3; c++) sum[c] += (double)vel[t->sc_list[j] * st + c];
    int64_t den = (int64_t)t->sc_cnt[k] + (accum ? 1 : 0);
    for (int tz = 0; tz < sc; tz++)
      for (int ty = 0; ty < sc; ty++)
        for (int tx = 0; tx < sc; tx++) {
          int32_t v[3] = {t->sc_pos[3 * k] + tx, t->sc_pos[3 * k + 1] + ty, t->sc_pos[3 * k + 2] + tz};
          int64_t ai = atlas_index(t, slot, v[0] - nd->pos[0] + a, v[1] - nd->pos[1] + a, v[2] - nd->pos[2] + a);
          for (int c = 0; c < 3; c++) {
            double* ch = t->chan[chans[c]];
            double s = sum[c] + (accum ? ch[ai] : 0.0);
            ch[ai] = den > 0 ? s / (double)den : 0.0;
          }
        }
  }
  return ORC_OK;
}

/* Colour part of gvdbGatherDensity / gvdbGatherLevelSet (P.cuh:713-731, 752-770): per voxel of
 * subcell k and component q of the uchar4 words,
 *   clr = (accum ? CHAR2CLR(old) : (0, 0, 0, 1)) + sum_j CHAR2CLR(clr_j);  clr /= cnt + accum;
 *   out = CLR2CHAR(clr) = uchar(clr * 255).
 * CHAR2CLR divides by 255 and CLR2CHAR multiplies by 255, so in units of 1/255 this is
 *   out_q = uchar((start_q + sum_j c_jq) / (cnt + accum)),  start = old or (0, 0, 0, 255).
 * Restated exactly: integer sums, the quotient floored (uchar truncates), saturated at 255
 * (the alpha sum exceeds cnt * 255 by the initial 255), and 0 for an empty subcell without
 * accumulation (the reference divides by zero there).  Channels hold the packed word
 * r | g << 8 | b << 16 | a << 24 as a double.  clr: uint32 words, stride in words. */
static uint32_t colour_mean(const uint64_t sum[4], uint64_t den, uint32_t start) {
  uint32_t out = 0;
  for (int q = 0; q < 4; q++) {
    uint64_t num = sum[q] + ((start >> (8 * q)) & 255u);
    uint64_t v = den ? num / den : 0;
    out |= (uint32_t)(v < 255 ? v : 255) << (8 * q);
  }
  return out;
}

int orc_gather_colour(orc_tree* t, const uint32_t* clr, int64_t st, int chan, int accum) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  if (!t->sc_cnt) return ORC_ERR_STATE;
  int sc = t->sc, a = t->apron;
  double* ch = t->chan[chan];
  for (int64_t k = 0; k < t->nsc; k++) {
    int64_t leaf = t->sc_nid[k];
    const orc_node* nd = &t->nodes[0][leaf];
    int64_t slot = t->leaf_slot[leaf];
    if (slot < 0) continue;
    uint64_t sum[4] = {0, 0, 0, 0};
    for (int64_t j = t->sc_off[k]; j < (int64_t)t->sc_off[k] + t->sc_cnt[k]; j++) {
      uint32_t w = clr[t->sc_list[j] * st];
      for (int q = 0; q < 4; q++) sum[q] += (w >> (8 * q)) & 255u;
    }
    uint64_t den = (uint64_t)t->sc_cnt[k] + (accum ? 1 : 0);
    for (int tz = 0; tz < sc; tz++)
      for (int ty = 0; ty < sc; ty++)
        for (int tx = 0; tx < sc; tx++) {
          int32_t v[3] = {t->sc_pos[3 * k] + tx, t->sc_pos[3 * k + 1] + ty, t->sc_pos[3 * k + 2] + tz};
          int64_t ai = atlas_index(t, slot, v[0] - nd->pos[0] + a, v[1] - nd->pos[1] + a, v[2] - nd->pos[2] + a);
          uint32_t start = accum ? (uint32_t)ch[ai] : 0xFF000000u;
          ch[ai] = (double)colour_mean(sum, den, start);
        }
  }
  return ORC_OK;
}

/* UpdateApron<T> (O.cuh:72-126): every apron voxel of every active brick takes the value
 * of the interior voxel at the same world position, or boundval if no leaf holds it. */
int orc_update_apron(orc_tree* t, int chan, double boundval) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  int a = t->apron;
  if (a == 0) return ORC_OK;
  int res0 = res_of(t, 0), br = res0 + 2 * a;
  double* ch = t->chan[chan];
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
  for (int64_t i = 0; i < t->cnt[0]; i++) {
    const orc_node* nd = &t->nodes[0][i];
    int64_t slot = t->leaf_slot[i];
    if (!nd->flags || slot < 0) continue;
    for (int z = 0; z < br; z++)
      for (int y = 0; y < br; y++)
        for (int x = 0; x < br; x++) {
          int inner = (x >= a && x < a + res0 && y >= a && y < a + res0 && z >= a && z < a + res0);
          if (inner) continue;
          int32_t g[3] = {nd->pos[0] + x - a, nd->pos[1] + y - a, nd->pos[2] + z - a};
          int64_t f = descend(t, g, 0);
          double v = boundval;
          if (f >= 0 && t->leaf_slot[f] >= 0) {
            const orc_node* nb = &t->nodes[0][f];
            v = ch[atlas_index(t, t->leaf_slot[f], g[0] - nb->pos[0] + a, g[1] - nb->pos[1] + a, g[2] - nb->pos[2] + a)];
          }
          ch[atlas_index(t, slot, x, y, z)] = v;
        }
  }
  return ORC_OK;
}

/* ScatterReduceLevelSet (new entry point): the GatherLevelSet field, produced by
 * scattering every (point, subcell) pair's |p-v|-r over the subcell's voxels and
 * min-reducing per voxel.  Identical values to orc_gather_levelset with init. */
int orc_scatter_reduce_levelset(orc_tree* t, float r, int chan, double init) {
  if (!t->chan[chan] || !t->sc_cnt) return ORC_ERR_STATE;
  int sc = t->sc, a = t->apron;
  double* ch = t->chan[chan];
  for (int64_t k = 0; k < t->nsc; k++) {
    int64_t leaf = t->sc_nid[k];
    const orc_node* nd = &t->nodes[0][leaf];
    int64_t slot = t->leaf_slot[leaf];
    if (slot < 0) continue;
    for (int tz = 0; tz < sc; tz++)
      for (int ty = 0; ty < sc; ty++)
        for (int tx = 0; tx < sc; tx++) {
          int32_t v[3] = {t->sc_pos[3 * k] + tx, t->sc_pos[3 * k + 1] + ty, t->sc_pos[3 * k + 2] + tz};
          ch[atlas_index(t, slot, v[0] - nd->pos[0] + a, v[1] - nd->pos[1] + a, v[2] - nd->pos[2] + a)] = init;
        }
  }
  for (int64_t k = 0; k < t->nsc; k++) {
    int64_t leaf = t->sc_nid[k];
    const orc_node* nd = &t->nodes[0][leaf];
    int64_t slot = t->leaf_slot[leaf];
    if (slot < 0) continue;
    for (int64_t j = t->sc_off[k]; j < (int64_t)t->sc_off[k] + t->sc_cnt[k]; j++)
      for (int tz = 0; tz < sc; tz++)
        for (int ty = 0; ty < sc; ty++)
          for (int tx = 0; tx < sc; tx++) {
            int32_t v[3] = {t->sc_pos[3 * k] + tx, t->sc_pos[3 * k + 1] + ty, t->sc_pos[3 * k + 2] + tz};
            double dx = (double)t->sc_list_pos[3 * j] - v[0];
            double dy = (double)t->sc_list_pos[3 * j + 1] - v[1];
            double dz = (double)t->sc_list_pos[3 * j + 2] - v[2];
            double w = sqrt(dx * dx + dy * dy + dz * dz) - (double)r;
            int64_t ai = atlas_index(t, slot, v[0] - nd->pos[0] + a, v[1] - nd->pos[1] + a, v[2] - nd->pos[2] + a);
            if (w < ch[ai]) ch[ai] = w;
          }
  }
  return ORC_OK;
}

/* ------------------------------------------------------------- trilinear */

/* Leaf holding world point q: getNodeAtPoint (N.cuh:243-253) on floor(q). */
static int64_t leaf_at(const orc_tree* t, const float* q) {
  int32_t p[3];
  for (int a = 0; a < 3; a++) {
    if (!(q[a] >= 0.0f)) return -1;
    p[a] = (int32_t)floorf(q[a]);
  }
  return descend(t, p, 0);
}

/* getTrilinear / getVolSampleF (R.cuh:104-126): tex3D linear filtering with voxel v
 * centred at v+0.5, read from the leaf's padded brick (apron included); 0 if no leaf. */
int orc_sample_trilinear(const orc_tree* t, int chan, const float* q, int64_t n, double* out) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  int a = t->apron;
  const double* ch = t->chan[chan];
  for (int64_t i = 0; i < n; i++) {
    int64_t f = leaf_at(t, q + 3 * i);
    if (f < 0 || t->leaf_slot[f] < 0) { out[i] = 0.0; continue; }
    const orc_node* nd = &t->nodes[0][f];
    int i0[3]; double al[3];
    for (int c = 0; c < 3; c++) {
      double x = (double)q[3 * i + c] - nd->pos[c] + a - 0.5;
      double fl = floor(x);
      i0[c] = (int)fl;
      al[c] = x - fl;
    }
    double s = 0;
    for (int dz = 0; dz < 2; dz++)
      for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
          double w = (dx ? al[0] : 1 - al[0]) * (dy ? al[1] : 1 - al[1]) * (dz ? al[2] : 1 - al[2]);
          s += w * ch[atlas_index(t, t->leaf_slot[f], i0[0] + dx, i0[1] + dy, i0[2] + dz)];
        }
    out[i] = s;
  }
  return ORC_OK;
}

/* Global-lattice trilinear corner weights of point q (voxel g centred at g+0.5). */
static void tri_corners(const float* q, int32_t* g0, double* al) {
  for (int c = 0; c < 3; c++) {
    double x = (double)q[c] - 0.5;
    double fl = floor(x);
    g0[c] = (int32_t)fl;
    al[c] = x - fl;
  }
}

/* Backward of sampling with respect to the query points: dq_i = dout_i * grad sample(q_i),
 * the derivative of the same padded-brick trilinear filter (0 where the sample is 0).  The
 * backward with respect to the field is the splat of dout (orc_splat_trilinear). */
int orc_sample_backward(const orc_tree* t, int chan, const float* q, const double* dout, int64_t n, double* dq) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  int a = t->apron;
  const double* ch = t->chan[chan];
  for (int64_t i = 0; i < n; i++) {
    dq[3 * i] = dq[3 * i + 1] = dq[3 * i + 2] = 0;
    int64_t f = leaf_at(t, q + 3 * i);
    if (f < 0 || t->leaf_slot[f] < 0) continue;
    const orc_node* nd = &t->nodes[0][f];
    int i0[3]; double al[3];
    for (int c = 0; c < 3; c++) {
      double x = (double)q[3 * i + c] - nd->pos[c] + a - 0.5;
      double fl = floor(x);
      i0[c] = (int)fl;
      al[c] = x - fl;
    }
    double gx = 0, gy = 0, gz = 0;
    for (int dz = 0; dz < 2; dz++)
      for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
          double v = ch[atlas_index(t, t->leaf_slot[f], i0[0] + dx, i0[1] + dy, i0[2] + dz)];
          double wx = dx ? al[0] : 1 - al[0], wy = dy ? al[1] : 1 - al[1], wz = dz ? al[2] : 1 - al[2];
          gx += (dx ? 1.0 : -1.0) * wy * wz * v;
          gy += wx * (dy ? 1.0 : -1.0) * wz * v;
          gz += wx * wy * (dz ? 1.0 : -1.0) * v;
        }
    dq[3 * i] = dout[i] * gx; dq[3 * i + 1] = dout[i] * gy; dq[3 * i + 2] = dout[i] * gz;
  }
  return ORC_OK;
}

/* ScatterDensity (gvdbScatterPointDensity, P.cuh:516-573, with the leaf of q = p + trans found
 * as gvdbInsertPoints does, P.cuh:19-36): for a valid point in a leaf with corner b, p' = q - b
 * (FP32), pi = trunc(p'); the voxel pi, and with expand its 6 face neighbours (unclamped, so
 * they may lie in the brick's apron; dropped when there is no apron), gains
 * distFunc = 1 + c(-3 + c(3 - c)) = (1 - c)^3 with c = |v - p'|^2 / r^2 evaluated in FP32 as the
 * reference does (P.cuh:514-521).  The reference's read-modify-write is racy; here every
 * contribution is summed (FP64). */
static float scatter_weight(const float* p, float vx, float vy, float vz, float r) {
  float bx = vx - p[0], by = vy - p[1], bz = vz - p[2];
  float c = (bx * bx + by * by + bz * bz) / (r * r);
  return (float)(1.0 + (double)(c * (-3.0f + c * (3.0f - c))));
}

int orc_scatter_density(orc_tree* t, int chan, const float* pos, int64_t n, int64_t stride, float r,
                        const float* trans, int expand) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  if (!(r > 0.0f)) return ORC_ERR_ARG;
  const int a = t->apron, res = res_of(t, 0);
  double* ch = t->chan[chan];
  for (int64_t i = 0; i < n; i++) {
    const float* pp = (const float*)((const char*)pos + i * stride);
    float q[3];
    for (int k = 0; k < 3; k++) q[k] = pp[k] + (trans ? trans[k] : 0.0f);
    if (!(q[0] < 1073741824.0f && q[1] < 1073741824.0f && q[2] < 1073741824.0f)) continue;  /* NOHIT etc. */
    int64_t f = leaf_at(t, q);
    if (f < 0 || t->leaf_slot[f] < 0) continue;
    const orc_node* nd = &t->nodes[0][f];
    float pl[3];
    int pi[3];
    for (int k = 0; k < 3; k++) {
      pl[k] = q[k] - (float)nd->pos[k];
      pi[k] = (int)pl[k];
    }
    if (pi[0] < 0 || pi[1] < 0 || pi[2] < 0 || pi[0] >= res || pi[1] >= res || pi[2] >= res) continue;
    static const int off[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int k = 0; k < (expand ? 7 : 1); k++) {
      const int v[3] = {pi[0] + off[k][0], pi[1] + off[k][1], pi[2] + off[k][2]};
      if (v[0] < -a || v[1] < -a || v[2] < -a || v[0] >= res + a || v[1] >= res + a || v[2] >= res + a) continue;
      ch[atlas_index(t, t->leaf_slot[f], v[0] + a, v[1] + a, v[2] + a)] +=
          (double)scatter_weight(pl, (float)v[0], (float)v[1], (float)v[2], r);
    }
  }
  return ORC_OK;
}

/* Colour part of gvdbScatterPointDensity (P.cuh:553-572) and gvdbScatterPointAvgCol
 * (P.cuh:610-641): for every point that reaches the density part with its voxel pi inside the
 * brick, avg != 0 counts it and sums its r, g, b at pi, then every voxel with a count takes
 * (r / count, g / count, b / count, 255) (unsigned integer division, as the reference);
 * avg == 0 gives the voxel the colour of one of its points -- the reference's unordered last
 * write, defined as the point with the largest index.  Voxels without points keep their value.
 * Channel values are the packed words as doubles; clr: one uint32 word per point. */
int orc_scatter_colour(orc_tree* t, int chan, const float* pos, int64_t n, int64_t stride, const float* trans,
                       const uint32_t* clr, int avg) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  const int a = t->apron, res = res_of(t, 0);
  const int64_t nslots = t->nslots, nin = nslots * res * res * res;
  uint64_t* sums = (uint64_t*)calloc((size_t)(nin > 0 ? nin : 1) * 4, sizeof(uint64_t));
  int64_t* last = (int64_t*)malloc((size_t)(nin > 0 ? nin : 1) * sizeof(int64_t));
  if (!sums || !last) { free(sums); free(last); return ORC_ERR_NOMEM; }
  for (int64_t v = 0; v < nin; v++) last[v] = -1;
  for (int64_t i = 0; i < n; i++) {
    const float* pp = (const float*)((const char*)pos + i * stride);
    float q[3];
    for (int k = 0; k < 3; k++) q[k] = pp[k] + (trans ? trans[k] : 0.0f);
    if (!(q[0] < 1073741824.0f && q[1] < 1073741824.0f && q[2] < 1073741824.0f)) continue;
    int64_t f = leaf_at(t, q);
    if (f < 0 || t->leaf_slot[f] < 0) continue;
    const orc_node* nd = &t->nodes[0][f];
    int pi[3];
    for (int k = 0; k < 3; k++) pi[k] = (int)(q[k] - (float)nd->pos[k]);
    if (pi[0] < 0 || pi[1] < 0 || pi[2] < 0 || pi[0] >= res || pi[1] >= res || pi[2] >= res) continue;
    const int64_t v = ((t->leaf_slot[f] * res + pi[2]) * res + pi[1]) * res + pi[0];
    const uint32_t w = clr[i];
    sums[4 * v] += 1;
    for (int c = 0; c < 3; c++) sums[4 * v + 1 + c] += (w >> (8 * c)) & 255u;
    last[v] = i;  /* increasing i: the largest index stays */
  }
  double* ch = t->chan[chan];
  for (int64_t v = 0; v < nin; v++) {
    const uint64_t cnt = sums[4 * v];
    if (!cnt) continue;
    uint32_t word;
    if (avg)
      word = (uint32_t)(sums[4 * v + 1] / cnt) | (uint32_t)(sums[4 * v + 2] / cnt) << 8 |
             (uint32_t)(sums[4 * v + 3] / cnt) << 16 | 255u << 24;
    else
      word = clr[last[v]];
    const int64_t slot = v / ((int64_t)res * res * res);
    const int l = (int)(v % ((int64_t)res * res * res));
    ch[atlas_index(t, slot, l % res + a, (l / res) % res + a, l / (res * res) + a)] = (double)word;
  }
  free(sums);
  free(last);
  return ORC_OK;
}

/* Trilinear splat rho(g) += m * w_g(q) into interior voxels of the leaves holding the
 * corners (the transpose of sampling over the global voxel lattice). */
int orc_splat_trilinear(orc_tree* t, int chan, const float* q, const float* mass, int64_t n) {
  if (!t->chan[chan]) return ORC_ERR_STATE;
  int a = t->apron;
  double* ch = t->chan[chan];
  for (int64_t i = 0; i < n; i++) {
    if (q[3 * i] < 0 || q[3 * i + 1] < 0 || q[3 * i + 2] < 0) continue;
    if (leaf_at(t, q + 3 * i) < 0) continue;
    int32_t g0[3]; double al[3];
    tri_corners(q + 3 * i, g0, al);
    double m = mass ? (double)mass[i] : 1.0;
    for (int dz = 0; dz < 2; dz++)
      for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
          int32_t g[3] = {g0[0] + dx, g0[1] + dy, g0[2] + dz};
          int64_t f = descend(t, g, 0);
          if (f < 0 || t->leaf_slot[f] < 0) continue;
          const orc_node* nb = &t->nodes[0][f];
          double w = (dx ? al[0] : 1 - al[0]) * (dy ? al[1] : 1 - al[1]) * (dz ? al[2] : 1 - al[2]);
          ch[atlas_index(t, t->leaf_slot[f], g[0] - nb->pos[0] + a, g[1] - nb->pos[1] + a, g[2] - nb->pos[2] + a)] +=
              m * w;
        }
  }
  return ORC_OK;
}

/* Backward of the splat: given dL/drho in chan_grad (interior voxels), dL/dm_i =
 * sum_g w_g(q_i) G(g) and dL/dq_i = m_i sum_g grad w_g(q_i) G(g). */
int orc_splat_backward(const orc_tree* t, int chan_grad, const float* q, const float* mass, int64_t n, double* dq,
                       double* dm) {
  if (!t->chan[chan_grad]) return ORC_ERR_STATE;
  int a = t->apron;
  const double* ch = t->chan[chan_grad];
  for (int64_t i = 0; i < n; i++) {
    dq[3 * i] = dq[3 * i + 1] = dq[3 * i + 2] = 0;
    dm[i] = 0;
    if (q[3 * i] < 0 || q[3 * i + 1] < 0 || q[3 * i + 2] < 0) continue;
    if (leaf_at(t, q + 3 * i) < 0) continue;
    int32_t g0[3]; double al[3];
    tri_corners(q + 3 * i, g0, al);
    double m = mass ? (double)mass[i] : 1.0;
    double s = 0, gx = 0, gy = 0, gz = 0;
    for (int dz = 0; dz < 2; dz++)
      for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
          int32_t g[3] = {g0[0] + dx, g0[1] + dy, g0[2] + dz};
          int64_t f = descend(t, g, 0);
          if (f < 0 || t->leaf_slot[f] < 0) continue;
          const orc_node* nb = &t->nodes[0][f];
          double G = ch[atlas_index(t, t->leaf_slot[f], g[0] - nb->pos[0] + a, g[1] - nb->pos[1] + a, g[2] - nb->pos[2] + a)];
          double wx = dx ? al[0] : 1 - al[0], wy = dy ? al[1] : 1 - al[1], wz = dz ? al[2] : 1 - al[2];
          double sx = dx ? 1.0 : -1.0, sy = dy ? 1.0 : -1.0, sz = dz ? 1.0 : -1.0;
          s += wx * wy * wz * G;
          gx += sx * wy * wz * G;
          gy += wx * sy * wz * G;
          gz += wx * wy * sz * G;
        }
    dm[i] = s;
    dq[3 * i] = m * gx; dq[3 * i + 1] = m * gy; dq[3 * i + 2] = m * gz;
  }
  return ORC_OK;
}

/* ------------------------------------------------------------ CPU baseline */

/* Per-point recursive activation from a root covering the bounding box: the host
 * ActivateSpace path (V.cpp:1220-1227, 2766-2855, AddChildNode V.cpp:2707-2723),
 * restated without its reparenting (the root is created at the covering depth). */
int64_t orc_cpu_activate_space(int levs, const int* ld, const float* pos, int64_t n, int64_t st, const float* tr,
                               int64_t* leaf_count) {
  orc_tree* t = orc_create(levs, ld, 16, 16, 1);
  if (!t) return -1;
  bounding_box(t, pos, n, st, tr);
  int32_t root[3];
  int d = determine_depth(t, root);
  if (check_depth(t, d)) { orc_destroy(t); return -1; }
  t->depth = d;
  memcpy(t->rootpos, root, sizeof(root));
  alloc_node(t, d, root[0], root[1], root[2]);
  for (int64_t i = 0; i < n; i++) {
    float q[3];
    for (int a = 0; a < 3; a++) q[a] = pos[i * st + a] + tr[a];
    if (q[0] < 0 || q[1] < 0 || q[2] < 0) continue;
    int32_t p[3] = {(int32_t)q[0], (int32_t)q[1], (int32_t)q[2]};
    int64_t cur = 0;
    for (int l = d; l > 0; l--) {
      orc_node* nd = &t->nodes[l][cur];
      int res = res_of(t, l);
      int64_t cr = range_of(t, l) / res;
      int64_t b[3];
      for (int a = 0; a < 3; a++) b[a] = ((int64_t)p[a] - nd->pos[a]) / cr;
      uint64_t ci = nd->childlist >> 16;
      uint64_t* e = &t->clist[l][ci * ((uint64_t)res * res * res) + (uint64_t)((b[2] * res + b[1]) * res + b[0])];
      if (*e == ORC_UNDEF64) {
        int32_t cp[3] = {(int32_t)(nd->pos[0] + b[0] * cr), (int32_t)(nd->pos[1] + b[1] * cr),
                         (int32_t)(nd->pos[2] + b[2] * cr)};
        int64_t c = alloc_node(t, l - 1, cp[0], cp[1], cp[2]);
        /* re-fetch: alloc may have moved the pools */
        e = &t->clist[l][ci * ((uint64_t)res * res * res) + (uint64_t)((b[2] * res + b[1]) * res + b[0])];
        *e = elem(0, l - 1, (uint64_t)c);
        t->nodes[l - 1][c].parent = elem(0, l, (uint64_t)cur);
      }
      cur = (int64_t)(*e >> 16);
    }
  }
  int64_t leaves = t->cnt[0];
  if (leaf_count) *leaf_count = leaves;
  orc_destroy(t);
  return leaves;
}
