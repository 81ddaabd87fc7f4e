/*
 * nrx_oracle.c -- CPU FP64 restatement of NetRAX's network likelihood path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product library never links or calls it.
 *
 * What it restates (reference file:line, paths relative to the NetRAX tree):
 *  - displayed-tree configurations, bit r set = second parent
 *      src/helper/ReticulationConfigHelper.cpp:188-203
 *  - dead nodes and the displayed-tree root
 *      src/helper/NetworkFunctions.cpp:228-279
 *  - per-node CLV variants of the DP, one per distinct alive subtree
 *      src/likelihood/ImprovedLoglikelihood.cpp:194-408
 *  - CLV update with per-site scaling (pll_update_partials_single, libpll;
 *    call sites ImprovedLoglikelihood.cpp:145,185): parent = prod over
 *    children of P_child * child, rescale by 2^256 when every entry of the
 *    site is below 2^-256, scaler = sum of child scalers + rescale
 *  - root log-likelihood (pll_compute_root_loglikelihood, libpll; call site
 *    ImprovedLoglikelihood.cpp:450): sum_s w_s [log sum_c w_c sum_i pi_i
 *    CLV[s,c,i] + scaler_s * log(2^-256)]
 *  - root variants / tree log-probabilities
 *      src/graph/ReticulationConfigSet.cpp:70-130, ImprovedLoglikelihood.cpp:410-486
 *  - AVERAGE / BEST mixture with the min_interesting_tree_logprob cut
 *      ImprovedLoglikelihood.cpp:521-604, NetraxOptions.hpp:108
 *  - branch-length derivatives mixed over displayed trees
 *      src/likelihood/LikelihoodDerivatives.cpp:30-188, 291-344
 *  - reticulation-probability derivative: no reference function; closed
 *    form d log P(t)/dp_r = [first]/p_r - [second]/(1-p_r).
 *
 * Model arithmetic is independent of the product engine on purpose:
 * P(t) = exp(Q r t) by scaling-and-squaring Taylor series (the engine uses
 * an eigendecomposition), P' = rQ P, P'' = (rQ)^2 P; discrete Gamma rates by
 * Yang's (1994) mean method with an incomplete-gamma series / continued
 * fraction and a bisection quantile.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_ERR_NEGINF 2
#define ORC_ERR_MEM 3
#define ORC_ERR_TOO_MANY 4

#define ORC_MAX_RETIC 16

static const double SCALE_THRESHOLD = 0x1p-256;
static const double SCALE_FACTOR = 0x1p256;

typedef struct {
  int32_t n_nodes, n_tips, root, n_edges, n_retic;
  const int32_t *edge_parent, *edge_child;
  const int32_t *retic_first_edge, *retic_second_edge;
} orc_net;

typedef struct {
  int32_t n_sites, states, ncat;
  const uint32_t *tip_masks; /* [n_tips][n_sites] */
  const uint32_t *weights;   /* [n_sites] */
  const double *subst_rates; /* [S(S-1)/2], upper triangle row-major */
  const double *freqs;       /* [S] */
  double alpha;              /* Gamma shape; ignored when ncat == 1 */
} orc_part;

static char g_err[256];
const char *orc_last_error(void) { return g_err; }

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------
 * Discrete Gamma (Yang 1994, mean of each equiprobable category)
 * ---------------------------------------------------------------------- */

/* Regularised lower incomplete gamma P(a, x): power series for x < a + 1,
 * Lentz continued fraction for the upper tail otherwise. */
static double lower_gamma_reg(double a, double x) {
  if (x <= 0.0) return 0.0;
  double lg = lgamma(a);
  if (x < a + 1.0) {
    double sum = 1.0 / a, term = sum, ap = a;
    for (int n = 0; n < 100000; ++n) {
      ap += 1.0;
      term *= x / ap;
      sum += term;
      if (fabs(term) < fabs(sum) * 1e-17) break;
    }
    return sum * exp(-x + a * log(x) - lg);
  }
  double b = x + 1.0 - a, c = 1.0 / 1e-300, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; ++i) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < 1e-300) d = 1e-300;
    c = b + an / c;
    if (fabs(c) < 1e-300) c = 1e-300;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - exp(-x + a * log(x) - lg) * h;
}

/* x with P(a, x) = p, by bisection on [0, hi] to full double resolution. */
static double gamma_quantile(double a, double p) {
  double lo = 0.0, hi = a + 10.0 * sqrt(a) + 50.0;
  while (lower_gamma_reg(a, hi) < p) hi *= 2.0;
  for (int it = 0; it < 2000; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (lower_gamma_reg(a, mid) < p)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

void orc_gamma_rates(double alpha, int ncat, double *rates) {
  if (ncat == 1) {
    rates[0] = 1.0;
    return;
  }
  /* category k spans Gamma(alpha, rate alpha) quantiles [k/K, (k+1)/K]; its
   * mean rate is K * [P(alpha+1, alpha x_{k+1}) - P(alpha+1, alpha x_k)] */
  double prev = 0.0;
  for (int k = 0; k < ncat; ++k) {
    double cur;
    if (k == ncat - 1) {
      cur = 1.0;
    } else {
      double x = gamma_quantile(alpha, (double)(k + 1) / ncat); /* = alpha*cut */
      cur = lower_gamma_reg(alpha + 1.0, x);
    }
    rates[k] = (cur - prev) * ncat;
    prev = cur;
  }
}

/* ------------------------------------------------------------------------
 * Substitution model and transition matrices
 * ---------------------------------------------------------------------- */

static void rate_matrix(int S, const double *subst, const double *freqs,
                        double *Q) {
  int k = 0;
  memset(Q, 0, sizeof(double) * S * S);
  for (int i = 0; i < S; ++i)
    for (int j = i + 1; j < S; ++j) {
      Q[i * S + j] = subst[k] * freqs[j];
      Q[j * S + i] = subst[k] * freqs[i];
      ++k;
    }
  double mu = 0.0;
  for (int i = 0; i < S; ++i) {
    double rs = 0.0;
    for (int j = 0; j < S; ++j)
      if (j != i) rs += Q[i * S + j];
    Q[i * S + i] = -rs;
    mu += freqs[i] * rs;
  }
  for (int i = 0; i < S * S; ++i) Q[i] /= mu;
}

static void matmul(int S, const double *A, const double *B, double *C) {
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j) {
      double s = 0.0;
      for (int k = 0; k < S; ++k) s += A[i * S + k] * B[k * S + j];
      C[i * S + j] = s;
    }
}

/* exp(A) by scaling and squaring with a Taylor series. */
static void expm_taylor(int S, const double *A, double *E) {
  double norm = 0.0;
  for (int j = 0; j < S; ++j) {
    double c = 0.0;
    for (int i = 0; i < S; ++i) c += fabs(A[i * S + j]);
    if (c > norm) norm = c;
  }
  int sq = 0;
  while (norm > 0.25) {
    norm *= 0.5;
    ++sq;
  }
  double scale = ldexp(1.0, -sq);
  double *As = malloc(sizeof(double) * S * S);
  double *term = malloc(sizeof(double) * S * S);
  double *tmp = malloc(sizeof(double) * S * S);
  for (int i = 0; i < S * S; ++i) As[i] = A[i] * scale;
  for (int i = 0; i < S * S; ++i) E[i] = term[i] = (i % (S + 1) == 0) ? 1.0 : 0.0;
  for (int n = 1; n <= 30; ++n) {
    matmul(S, term, As, tmp);
    double mx = 0.0;
    for (int i = 0; i < S * S; ++i) {
      term[i] = tmp[i] / n;
      E[i] += term[i];
      if (fabs(term[i]) > mx) mx = fabs(term[i]);
    }
    if (mx < 1e-20) break;
  }
  for (int s = 0; s < sq; ++s) {
    matmul(S, E, E, tmp);
    memcpy(E, tmp, sizeof(double) * S * S);
  }
  free(As);
  free(term);
  free(tmp);
}

/* P = exp(Q r t), dP = rQ P, d2P = (rQ)^2 P; Q normalised to mean rate 1. */
void orc_pmatrix(int S, const double *subst, const double *freqs, double r,
                 double t, double *P, double *dP, double *d2P) {
  double *Q = malloc(sizeof(double) * S * S);
  double *A = malloc(sizeof(double) * S * S);
  rate_matrix(S, subst, freqs, Q);
  for (int i = 0; i < S * S; ++i) A[i] = Q[i] * r * t;
  expm_taylor(S, A, P);
  for (int i = 0; i < S * S; ++i) Q[i] *= r;
  if (dP) matmul(S, Q, P, dP);
  if (d2P) {
    if (!dP) {
      double *tmp = malloc(sizeof(double) * S * S);
      matmul(S, Q, P, tmp);
      matmul(S, Q, tmp, d2P);
      free(tmp);
    } else {
      matmul(S, Q, dP, d2P);
    }
  }
  free(Q);
  free(A);
}

/* ------------------------------------------------------------------------
 * Displayed-tree configurations and CLV variants
 * ---------------------------------------------------------------------- */

typedef struct {
  int node;
  int nkids;
  int kid_node[2];
  int kid_edge[2];
  int kid_entry[2]; /* -1 for a tip child */
  const uint64_t *key;
} entry_t;

typedef struct {
  int N, E, R, W, T;
  int *kid_start, *kid_edge; /* CSR of child edges per node */
  int *topo;                 /* parents before children */
  uint8_t *active;           /* [T][E] */
  uint8_t *alive;            /* [T][N] */
  uint64_t *key;             /* [T][N][W] alive edges below the node */
  int *entry_of;             /* [T][N] variant index at node, -1 if none */
  int *dtroot;               /* [T] displayed-tree root */
  int n_entries;
  entry_t *entries;
  int *entry_order; /* children before parents */
  /* classes = root variants */
  int n_classes;
  int *class_of; /* [T] */
  int *class_rep;
  double *class_logprob;
  double *class_dlogprob; /* [n_classes][R] */
} dag_t;

static void dag_free(dag_t *d) {
  free(d->kid_start);
  free(d->kid_edge);
  free(d->topo);
  free(d->active);
  free(d->alive);
  free(d->key);
  free(d->entry_of);
  free(d->dtroot);
  free(d->entries);
  free(d->entry_order);
  free(d->class_of);
  free(d->class_rep);
  free(d->class_logprob);
  free(d->class_dlogprob);
  memset(d, 0, sizeof(*d));
}

static int dag_build(const orc_net *net, const double *retic_prob, dag_t *d) {
  memset(d, 0, sizeof(*d));
  if (net->n_retic > ORC_MAX_RETIC) {
    snprintf(g_err, sizeof g_err, "oracle supports at most %d reticulations",
             ORC_MAX_RETIC);
    return ORC_ERR_TOO_MANY;
  }
  int N = net->n_nodes, E = net->n_edges, R = net->n_retic;
  int W = (E + 63) / 64, T = 1 << R;
  d->N = N;
  d->E = E;
  d->R = R;
  d->W = W;
  d->T = T;
  d->kid_start = calloc(N + 1, sizeof(int));
  d->kid_edge = malloc(sizeof(int) * E);
  for (int e = 0; e < E; ++e) d->kid_start[net->edge_parent[e] + 1]++;
  for (int v = 0; v < N; ++v) d->kid_start[v + 1] += d->kid_start[v];
  int *fill = calloc(N, sizeof(int));
  for (int e = 0; e < E; ++e) {
    int p = net->edge_parent[e];
    d->kid_edge[d->kid_start[p] + fill[p]++] = e;
  }
  free(fill);
  /* topological order (Kahn) */
  int *indeg = calloc(N, sizeof(int));
  for (int e = 0; e < E; ++e) indeg[net->edge_child[e]]++;
  d->topo = malloc(sizeof(int) * N);
  int head = 0, tail = 0;
  for (int v = 0; v < N; ++v)
    if (indeg[v] == 0) d->topo[tail++] = v;
  while (head < tail) {
    int v = d->topo[head++];
    for (int k = d->kid_start[v]; k < d->kid_start[v + 1]; ++k) {
      int c = net->edge_child[d->kid_edge[k]];
      if (--indeg[c] == 0) d->topo[tail++] = c;
    }
  }
  free(indeg);
  if (tail != N || d->topo[0] != net->root) {
    snprintf(g_err, sizeof g_err, "network is not a rooted DAG");
    return ORC_ERR_ARG;
  }

  d->active = malloc((size_t)T * E);
  d->alive = malloc((size_t)T * N);
  d->key = calloc((size_t)T * N * W, sizeof(uint64_t));
  d->entry_of = malloc(sizeof(int) * (size_t)T * N);
  d->dtroot = malloc(sizeof(int) * T);
  d->entries = malloc(sizeof(entry_t) * (size_t)T * N);
  for (int t = 0; t < T; ++t) {
    uint8_t *act = d->active + (size_t)t * E;
    uint8_t *alv = d->alive + (size_t)t * N;
    uint64_t *key = d->key + (size_t)t * N * W;
    memset(act, 1, E);
    for (int r = 0; r < R; ++r) {
      if ((t >> r) & 1)
        act[net->retic_first_edge[r]] = 0;
      else
        act[net->retic_second_edge[r]] = 0;
    }
    for (int i = N - 1; i >= 0; --i) {
      int v = d->topo[i];
      if (v < net->n_tips) {
        alv[v] = 1;
        continue;
      }
      alv[v] = 0;
      for (int k = d->kid_start[v]; k < d->kid_start[v + 1]; ++k) {
        int e = d->kid_edge[k], c = net->edge_child[e];
        if (act[e] && alv[c]) {
          alv[v] = 1;
          key[(size_t)v * W + e / 64] |= 1ull << (e % 64);
          for (int w = 0; w < W; ++w) key[(size_t)v * W + w] |= key[(size_t)c * W + w];
        }
      }
    }
    /* displayed-tree root: walk down single alive children */
    int u = net->root;
    for (;;) {
      int na = 0, only = -1;
      for (int k = d->kid_start[u]; k < d->kid_start[u + 1]; ++k) {
        int e = d->kid_edge[k], c = net->edge_child[e];
        if (act[e] && alv[c]) {
          ++na;
          only = c;
        }
      }
      if (na == 1)
        u = only;
      else
        break;
    }
    d->dtroot[t] = u;
  }
  /* variants: distinct (node, key) over configurations */
  int ne = 0;
  for (int t = 0; t < T; ++t)
    for (int v = 0; v < N; ++v) d->entry_of[(size_t)t * N + v] = -1;
  for (int i = N - 1; i >= 0; --i) {
    int v = d->topo[i];
    if (v < net->n_tips) continue;
    int first = ne;
    for (int t = 0; t < T; ++t) {
      if (!d->alive[(size_t)t * N + v]) continue;
      const uint64_t *k = d->key + ((size_t)t * N + v) * W;
      int found = -1;
      for (int j = first; j < ne; ++j)
        if (!memcmp(d->entries[j].key, k, sizeof(uint64_t) * W)) {
          found = j;
          break;
        }
      if (found < 0) {
        entry_t *en = &d->entries[ne];
        en->node = v;
        en->key = k;
        en->nkids = 0;
        const uint8_t *act = d->active + (size_t)t * E;
        const uint8_t *alv = d->alive + (size_t)t * N;
        for (int kk = d->kid_start[v]; kk < d->kid_start[v + 1]; ++kk) {
          int e = d->kid_edge[kk], c = net->edge_child[e];
          if (act[e] && alv[c]) {
            en->kid_node[en->nkids] = c;
            en->kid_edge[en->nkids] = e;
            en->kid_entry[en->nkids] =
                c < net->n_tips ? -1 : d->entry_of[(size_t)t * N + c];
            en->nkids++;
          }
        }
        found = ne++;
      }
      d->entry_of[(size_t)t * N + v] = found;
    }
  }
  d->n_entries = ne;
  d->entry_order = malloc(sizeof(int) * (ne ? ne : 1));
  for (int j = 0; j < ne; ++j) d->entry_order[j] = j; /* built children first */

  /* classes: configurations with the same alive edge set at the root */
  d->class_of = malloc(sizeof(int) * T);
  d->class_rep = malloc(sizeof(int) * T);
  int nc = 0;
  for (int t = 0; t < T; ++t) {
    const uint64_t *k = d->key + ((size_t)t * N + net->root) * W;
    int found = -1;
    for (int c = 0; c < nc; ++c) {
      const uint64_t *k2 = d->key + ((size_t)d->class_rep[c] * N + net->root) * W;
      if (!memcmp(k, k2, sizeof(uint64_t) * W)) {
        found = c;
        break;
      }
    }
    if (found < 0) {
      d->class_rep[nc] = t;
      found = nc++;
    }
    d->class_of[t] = found;
  }
  d->n_classes = nc;
  d->class_logprob = malloc(sizeof(double) * nc);
  d->class_dlogprob = calloc((size_t)(nc > 0 ? nc : 1) * (size_t)(R > 0 ? R : 1), sizeof(double));
  for (int c = 0; c < nc; ++c) {
    /* log sum over member configurations of prod_r (p_r or 1-p_r) */
    double mx = -INFINITY;
    int cnt = 0;
    for (int t = 0; t < T; ++t) {
      if (d->class_of[t] != c) continue;
      double lp = 0.0;
      for (int r = 0; r < R; ++r)
        lp += ((t >> r) & 1) ? log(1.0 - retic_prob[r]) : log(retic_prob[r]);
      if (lp > mx) mx = lp;
      ++cnt;
    }
    double s = 0.0;
    double *dl = d->class_dlogprob + (size_t)c * R;
    for (int t = 0; t < T; ++t) {
      if (d->class_of[t] != c) continue;
      double lp = 0.0;
      for (int r = 0; r < R; ++r)
        lp += ((t >> r) & 1) ? log(1.0 - retic_prob[r]) : log(retic_prob[r]);
      double wgt = exp(lp - mx);
      s += wgt;
      for (int r = 0; r < R; ++r)
        dl[r] += wgt * (((t >> r) & 1) ? -1.0 / (1.0 - retic_prob[r])
                                       : 1.0 / retic_prob[r]);
    }
    for (int r = 0; r < R; ++r) dl[r] /= s;
    if (cnt == 1) {
      int t = d->class_rep[c];
      double lp = 0.0;
      for (int r = 0; r < R; ++r)
        lp += ((t >> r) & 1) ? log(1.0 - retic_prob[r]) : log(retic_prob[r]);
      d->class_logprob[c] = lp;
    } else {
      d->class_logprob[c] = mx + log(s);
    }
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------------
 * Per-partition CLVs
 * ---------------------------------------------------------------------- */

typedef struct {
  int S, C, n;
  double *rates;  /* [C] */
  double *P;      /* [E][C][S][S] */
  double *dP, *d2P;
  double **clv;   /* [n_entries] -> [n][C][S] */
  uint32_t **scl; /* [n_entries] -> [n] */
} part_state;

static void tip_vec(uint32_t mask, int S, double *v) {
  for (int i = 0; i < S; ++i) v[i] = (mask >> i) & 1u ? 1.0 : 0.0;
}

static void part_setup(const orc_net *net, const double *brlen, const orc_part *p,
                       int with_derivs, part_state *ps) {
  int S = p->states, C = p->ncat, E = net->n_edges;
  ps->S = S;
  ps->C = C;
  ps->n = p->n_sites;
  ps->rates = malloc(sizeof(double) * C);
  orc_gamma_rates(p->alpha, C, ps->rates);
  ps->P = malloc(sizeof(double) * (size_t)E * C * S * S);
  ps->dP = with_derivs ? malloc(sizeof(double) * (size_t)E * C * S * S) : NULL;
  ps->d2P = with_derivs ? malloc(sizeof(double) * (size_t)E * C * S * S) : NULL;
  for (int e = 0; e < E; ++e)
    for (int c = 0; c < C; ++c) {
      size_t off = ((size_t)e * C + c) * S * S;
      orc_pmatrix(S, p->subst_rates, p->freqs, ps->rates[c], brlen[e], ps->P + off,
                  with_derivs ? ps->dP + off : NULL, with_derivs ? ps->d2P + off : NULL);
    }
  ps->clv = NULL;
  ps->scl = NULL;
}

static void part_free(part_state *ps, int n_entries) {
  if (ps->clv)
    for (int j = 0; j < n_entries; ++j) {
      free(ps->clv[j]);
      free(ps->scl[j]);
    }
  free(ps->clv);
  free(ps->scl);
  free(ps->rates);
  free(ps->P);
  free(ps->dP);
  free(ps->d2P);
}

/* y[c][i] = sum_j M[c][i][j] x[c][j] for one site */
static inline void matvec_site(int S, int C, const double *M, const double *x,
                               double *y) {
  for (int c = 0; c < C; ++c)
    for (int i = 0; i < S; ++i) {
      double s = 0.0;
      const double *row = M + ((size_t)c * S + i) * S;
      for (int j = 0; j < S; ++j) s += row[j] * x[c * S + j];
      y[c * S + i] = s;
    }
}

static int part_forward(const orc_net *net, const dag_t *d, const orc_part *p,
                        part_state *ps) {
  int S = ps->S, C = ps->C, n = ps->n, CS = C * S;
  ps->clv = calloc(d->n_entries ? d->n_entries : 1, sizeof(double *));
  ps->scl = calloc(d->n_entries ? d->n_entries : 1, sizeof(uint32_t *));
  for (int j = 0; j < d->n_entries; ++j) {
    ps->clv[j] = malloc(sizeof(double) * (size_t)n * CS);
    ps->scl[j] = malloc(sizeof(uint32_t) * (size_t)n);
    if (!ps->clv[j] || !ps->scl[j]) {
      snprintf(g_err, sizeof g_err, "oracle out of memory");
      return ORC_ERR_MEM;
    }
  }
  for (int jj = 0; jj < d->n_entries; ++jj) {
    const entry_t *en = &d->entries[d->entry_order[jj]];
    double *out = ps->clv[d->entry_order[jj]];
    uint32_t *osc = ps->scl[d->entry_order[jj]];
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
      double x[20 * 16], y[20 * 16];
      double *o = out + (size_t)s * CS;
      uint32_t sc = 0;
      for (int q = 0; q < CS; ++q) o[q] = 1.0;
      for (int k = 0; k < en->nkids; ++k) {
        const double *src;
        if (en->kid_entry[k] < 0) {
          uint32_t m = p->tip_masks[(size_t)en->kid_node[k] * n + s];
          for (int c = 0; c < C; ++c) tip_vec(m, S, x + c * S);
          src = x;
        } else {
          src = ps->clv[en->kid_entry[k]] + (size_t)s * CS;
          sc += ps->scl[en->kid_entry[k]][s];
        }
        matvec_site(S, C, ps->P + (size_t)en->kid_edge[k] * C * S * S, src, y);
        for (int q = 0; q < CS; ++q) o[q] *= y[q];
      }
      double mx = 0.0;
      for (int q = 0; q < CS; ++q)
        if (o[q] > mx) mx = o[q];
      if (mx < SCALE_THRESHOLD) {
        for (int q = 0; q < CS; ++q) o[q] *= SCALE_FACTOR;
        sc += 1;
      }
      osc[s] = sc;
    }
  }
  return ORC_OK;
}

/* lnL of one displayed tree (class) for one partition */
static double part_root_logl(const dag_t *d, const orc_part *p, const part_state *ps,
                             int t) {
  int S = ps->S, C = ps->C, n = ps->n, CS = C * S;
  int v = d->dtroot[t];
  int j = d->entry_of[(size_t)t * d->N + v];
  const double *clv = ps->clv[j];
  const uint32_t *scl = ps->scl[j];
  double logthr = log(SCALE_THRESHOLD);
  double total = 0.0;
#pragma omp parallel for reduction(+ : total) schedule(static)
  for (int s = 0; s < n; ++s) {
    double L = 0.0;
    for (int c = 0; c < C; ++c) {
      double t2 = 0.0;
      for (int i = 0; i < S; ++i) t2 += p->freqs[i] * clv[(size_t)s * CS + c * S + i];
      L += t2 / C;
    }
    total += p->weights[s] * (log(L) + scl[s] * logthr);
  }
  return total;
}

/* Outside pass over the displayed tree of configuration t: per-edge
 * first and second derivatives of lnL_t (edges outside the tree get 0). */
static void part_tree_derivs(const orc_net *net, const dag_t *d, const orc_part *p,
                             const part_state *ps, int t, double *d1, double *d2) {
  int S = ps->S, C = ps->C, n = ps->n, CS = C * S, N = d->N, E = d->E;
  const uint8_t *act = d->active + (size_t)t * E;
  const uint8_t *alv = d->alive + (size_t)t * N;
  for (int e = 0; e < E; ++e) d1[e] = d2[e] = 0.0;
  /* nodes of the displayed tree below dtroot, parents before children */
  int *stack = malloc(sizeof(int) * N);
  int *order = malloc(sizeof(int) * N);
  int no = 0, sp = 0;
  stack[sp++] = d->dtroot[t];
  while (sp) {
    int v = stack[--sp];
    order[no++] = v;
    for (int k = d->kid_start[v]; k < d->kid_start[v + 1]; ++k) {
      int e = d->kid_edge[k], c = net->edge_child[e];
      if (act[e] && alv[c] && c >= net->n_tips) stack[sp++] = c;
    }
  }
  /* up[v] = outside vector at node v (excluding v's own subtree) */
  double **up = calloc(N, sizeof(double *));
  for (int i = 0; i < no; ++i) up[order[i]] = malloc(sizeof(double) * (size_t)n * CS);
  {
    double *u = up[d->dtroot[t]];
    for (int s = 0; s < n; ++s)
      for (int c = 0; c < C; ++c)
        for (int i = 0; i < S; ++i) u[(size_t)s * CS + c * S + i] = p->freqs[i];
  }
  for (int oi = 0; oi < no; ++oi) {
    int u = order[oi];
    int kids[2], kedges[2], nk = 0;
    for (int k = d->kid_start[u]; k < d->kid_start[u + 1]; ++k) {
      int e = d->kid_edge[k], c = net->edge_child[e];
      if (act[e] && alv[c]) {
        kids[nk] = c;
        kedges[nk] = e;
        ++nk;
      }
    }
    for (int a = 0; a < nk; ++a) {
      int v = kids[a], e = kedges[a];
      int jv = v < net->n_tips ? -1 : d->entry_of[(size_t)t * N + v];
      double g1 = 0.0, g2 = 0.0;
#pragma omp parallel for reduction(+ : g1, g2) schedule(static)
      for (int s = 0; s < n; ++s) {
        double out[20 * 16], x[20 * 16], y[20 * 16], tmp[20 * 16];
        for (int q = 0; q < CS; ++q) out[q] = up[u][(size_t)s * CS + q];
        for (int b = 0; b < nk; ++b) {
          if (b == a) continue;
          int w = kids[b];
          const double *src;
          if (w < net->n_tips) {
            uint32_t m = p->tip_masks[(size_t)w * n + s];
            for (int c = 0; c < C; ++c) tip_vec(m, S, x + c * S);
            src = x;
          } else {
            src = ps->clv[d->entry_of[(size_t)t * N + w]] + (size_t)s * CS;
          }
          matvec_site(S, C, ps->P + (size_t)kedges[b] * C * S * S, src, y);
          for (int q = 0; q < CS; ++q) out[q] *= y[q];
        }
        const double *dn;
        if (jv < 0) {
          uint32_t m = p->tip_masks[(size_t)v * n + s];
          for (int c = 0; c < C; ++c) tip_vec(m, S, x + c * S);
          dn = x;
        } else {
          dn = ps->clv[jv] + (size_t)s * CS;
        }
        double den = 0.0, n1 = 0.0, n2 = 0.0;
        size_t off = (size_t)e * C * S * S;
        matvec_site(S, C, ps->P + off, dn, y);
        for (int q = 0; q < CS; ++q) den += out[q] * y[q];
        matvec_site(S, C, ps->dP + off, dn, y);
        for (int q = 0; q < CS; ++q) n1 += out[q] * y[q];
        matvec_site(S, C, ps->d2P + off, dn, y);
        for (int q = 0; q < CS; ++q) n2 += out[q] * y[q];
        double r1 = n1 / den, r2 = n2 / den;
        g1 += p->weights[s] * r1;
        g2 += p->weights[s] * (r2 - r1 * r1);
        if (jv >= 0) {
          /* up[v][c][j] = sum_i out[c][i] P[c][i][j], rescaled when tiny */
          double *uv = up[v] + (size_t)s * CS;
          double mx = 0.0;
          for (int c = 0; c < C; ++c)
            for (int jj = 0; jj < S; ++jj) {
              double acc = 0.0;
              for (int i = 0; i < S; ++i)
                acc += out[c * S + i] * ps->P[off + ((size_t)c * S + i) * S + jj];
              tmp[c * S + jj] = acc;
              if (acc > mx) mx = acc;
            }
          double f = mx < SCALE_THRESHOLD ? SCALE_FACTOR : 1.0;
          for (int q = 0; q < CS; ++q) uv[q] = tmp[q] * f;
        }
      }
      d1[e] = g1;
      d2[e] = g2;
    }
  }
  for (int i = 0; i < N; ++i) free(up[i]);
  free(up);
  free(stack);
  free(order);
}

/* ------------------------------------------------------------------------
 * Public entry points
 * ---------------------------------------------------------------------- */

static int check_args(const orc_net *net, const orc_part *parts, int n_parts) {
  if (!net || !parts || n_parts <= 0) {
    snprintf(g_err, sizeof g_err, "bad arguments");
    return ORC_ERR_ARG;
  }
  for (int p = 0; p < n_parts; ++p)
    if (parts[p].states > 20 || parts[p].ncat > 16 || parts[p].n_sites <= 0) {
      snprintf(g_err, sizeof g_err, "unsupported partition shape");
      return ORC_ERR_ARG;
    }
  return ORC_OK;
}

/* Shared driver: forward over all partitions, per (partition, class) lnL,
 * mixture.  Optionally returns per-class derivative data. */
typedef struct {
  int n_parts, n_classes;
  double *tree_logl; /* [P][n_classes] */
  int *use;          /* [n_classes] passes the logprob cut */
} fwd_result;

int orc_num_classes(const orc_net *net, const double *retic_prob) {
  dag_t d;
  int rc = dag_build(net, retic_prob, &d);
  if (rc) return -1;
  int nc = d.n_classes;
  dag_free(&d);
  return nc;
}

int orc_num_variants(const orc_net *net, const double *retic_prob) {
  dag_t d;
  int rc = dag_build(net, retic_prob, &d);
  if (rc) return -1;
  int ne = d.n_entries;
  dag_free(&d);
  return ne;
}

static double mixture(int variant, int nc, const double *logprob, const int *use,
                      const double *logl, double *alpha) {
  double best = -INFINITY;
  int arg = -1;
  for (int c = 0; c < nc; ++c) {
    if (!use[c]) continue;
    double v = logprob[c] + logl[c];
    if (v > best) {
      best = v;
      arg = c;
    }
  }
  if (variant == 1) {
    for (int c = 0; c < nc; ++c) alpha[c] = (c == arg) ? 1.0 : 0.0;
    return best;
  }
  double s = 0.0;
  for (int c = 0; c < nc; ++c) {
    alpha[c] = use[c] ? exp(logprob[c] + logl[c] - best) : 0.0;
    s += alpha[c];
  }
  for (int c = 0; c < nc; ++c) alpha[c] /= s;
  return best + log(s);
}

/* variant: 0 = AVERAGE_DISPLAYED_TREES, 1 = BEST_DISPLAYED_TREE
 * (src/likelihood/LikelihoodVariant.hpp:3-7).  out_tree_logl, when given,
 * receives [n_parts][n_classes]; out_class_logprob [n_classes]. */
int orc_loglikelihood(const orc_net *net, const double *brlen, const double *retic_prob,
                      const orc_part *parts, int n_parts, int variant,
                      double min_tree_logprob, double *out_total, double *out_part,
                      double *out_tree_logl, double *out_class_logprob) {
  int rc = check_args(net, parts, n_parts);
  if (rc) return rc;
  dag_t d;
  rc = dag_build(net, retic_prob, &d);
  if (rc) return rc;
  int nc = d.n_classes;
  int *use = malloc(sizeof(int) * nc);
  double *logl = malloc(sizeof(double) * nc);
  double *alpha = malloc(sizeof(double) * nc);
  for (int c = 0; c < nc; ++c) use[c] = d.class_logprob[c] >= min_tree_logprob;
  double total = 0.0;
  for (int p = 0; p < n_parts && rc == ORC_OK; ++p) {
    part_state ps;
    part_setup(net, brlen, &parts[p], 0, &ps);
    rc = part_forward(net, &d, &parts[p], &ps);
    if (rc == ORC_OK) {
      for (int c = 0; c < nc; ++c) {
        logl[c] = use[c] || out_tree_logl
                      ? part_root_logl(&d, &parts[p], &ps, d.class_rep[c])
                      : 0.0;
        if (out_tree_logl) out_tree_logl[(size_t)p * nc + c] = logl[c];
      }
      double lp = mixture(variant, nc, d.class_logprob, use, logl, alpha);
      if (out_part) out_part[p] = lp;
      total += lp;
    }
    part_free(&ps, d.n_entries);
  }
  if (out_class_logprob)
    for (int c = 0; c < nc; ++c) out_class_logprob[c] = d.class_logprob[c];
  free(use);
  free(logl);
  free(alpha);
  dag_free(&d);
  if (rc) return rc;
  if (total == -INFINITY || isnan(total)) {
    snprintf(g_err, sizeof g_err, "Invalid network likelihood: negative infinity");
    return ORC_ERR_NEGINF;
  }
  *out_total = total;
  return ORC_OK;
}

/* Full gradient of the mixture lnL: d/d(branch length) for every edge and
 * d/d(first-parent probability) for every reticulation.  out_d2 (optional)
 * receives the diagonal second derivative per edge. */
int orc_gradient(const orc_net *net, const double *brlen, const double *retic_prob,
                 const orc_part *parts, int n_parts, int variant, double min_tree_logprob,
                 double *out_total, double *out_dbrlen, double *out_dprob,
                 double *out_part_dbrlen /* [P][E] or NULL */) {
  int rc = check_args(net, parts, n_parts);
  if (rc) return rc;
  dag_t d;
  rc = dag_build(net, retic_prob, &d);
  if (rc) return rc;
  int nc = d.n_classes, E = net->n_edges, R = net->n_retic;
  int *use = malloc(sizeof(int) * nc);
  double *logl = malloc(sizeof(double) * nc);
  double *alpha = malloc(sizeof(double) * nc);
  double *t1 = malloc(sizeof(double) * E), *t2 = malloc(sizeof(double) * E);
  for (int c = 0; c < nc; ++c) use[c] = d.class_logprob[c] >= min_tree_logprob;
  for (int e = 0; e < E; ++e) out_dbrlen[e] = 0.0;
  for (int r = 0; r < R; ++r) out_dprob[r] = 0.0;
  double total = 0.0;
  for (int p = 0; p < n_parts && rc == ORC_OK; ++p) {
    part_state ps;
    part_setup(net, brlen, &parts[p], 1, &ps);
    rc = part_forward(net, &d, &parts[p], &ps);
    if (rc == ORC_OK) {
      for (int c = 0; c < nc; ++c)
        logl[c] = use[c] ? part_root_logl(&d, &parts[p], &ps, d.class_rep[c]) : 0.0;
      total += mixture(variant, nc, d.class_logprob, use, logl, alpha);
      if (out_part_dbrlen)
        for (int e = 0; e < E; ++e) out_part_dbrlen[(size_t)p * E + e] = 0.0;
      for (int c = 0; c < nc; ++c) {
        if (alpha[c] == 0.0) continue;
        part_tree_derivs(net, &d, &parts[p], &ps, d.class_rep[c], t1, t2);
        for (int e = 0; e < E; ++e) {
          out_dbrlen[e] += alpha[c] * t1[e];
          if (out_part_dbrlen) out_part_dbrlen[(size_t)p * E + e] += alpha[c] * t1[e];
        }
        for (int r = 0; r < R; ++r)
          out_dprob[r] += alpha[c] * d.class_dlogprob[(size_t)c * R + r];
      }
    }
    part_free(&ps, d.n_entries);
  }
  free(use);
  free(logl);
  free(alpha);
  free(t1);
  free(t2);
  dag_free(&d);
  if (rc) return rc;
  if (total == -INFINITY || isnan(total)) {
    snprintf(g_err, sizeof g_err, "Invalid network likelihood: negative infinity");
    return ORC_ERR_NEGINF;
  }
  *out_total = total;
  return ORC_OK;
}

/* First and second lnL derivatives for one branch, mixed over displayed
 * trees as computeLoglikelihoodDerivatives does
 * (LikelihoodDerivatives.cpp:30-232).
 *   mode 0 = reference normalisation: only trees in which the branch is
 *            active and alive enter the sums (LikelihoodDerivatives.cpp:325),
 *            BEST picks the tree maximising tree_logl * tree_prob (:161-167);
 *   mode 1 = true derivatives of the network lnL. */
int orc_brlen_derivatives(const orc_net *net, const double *brlen,
                          const double *retic_prob, const orc_part *parts, int n_parts,
                          int variant, double min_tree_logprob, int edge, int mode,
                          double *out_d1, double *out_d2, double *out_part_d1,
                          double *out_part_d2) {
  int rc = check_args(net, parts, n_parts);
  if (rc) return rc;
  if (edge < 0 || edge >= net->n_edges) {
    snprintf(g_err, sizeof g_err, "edge index out of range");
    return ORC_ERR_ARG;
  }
  dag_t d;
  rc = dag_build(net, retic_prob, &d);
  if (rc) return rc;
  int nc = d.n_classes, E = net->n_edges, N = d.N, W = d.W;
  int *use = malloc(sizeof(int) * nc);
  double *logl = malloc(sizeof(double) * nc);
  double *l1 = malloc(sizeof(double) * nc), *l2 = malloc(sizeof(double) * nc);
  double *alpha = malloc(sizeof(double) * nc);
  double *t1 = malloc(sizeof(double) * E), *t2 = malloc(sizeof(double) * E);
  int *has_edge = malloc(sizeof(int) * nc);
  for (int c = 0; c < nc; ++c) {
    use[c] = d.class_logprob[c] >= min_tree_logprob;
    const uint64_t *k = d.key + ((size_t)d.class_rep[c] * N + net->root) * W;
    has_edge[c] = (k[edge / 64] >> (edge % 64)) & 1ull;
  }
  double D1 = 0.0, D2 = 0.0;
  for (int p = 0; p < n_parts && rc == ORC_OK; ++p) {
    part_state ps;
    part_setup(net, brlen, &parts[p], 1, &ps);
    rc = part_forward(net, &d, &parts[p], &ps);
    double pd1 = 0.0, pd2 = 0.0;
    if (rc == ORC_OK) {
      for (int c = 0; c < nc; ++c) {
        int need = use[c] && (mode == 1 || has_edge[c]);
        logl[c] = l1[c] = l2[c] = 0.0;
        if (!need) continue;
        logl[c] = part_root_logl(&d, &parts[p], &ps, d.class_rep[c]);
        part_tree_derivs(net, &d, &parts[p], &ps, d.class_rep[c], t1, t2);
        l1[c] = t1[edge];
        l2[c] = t2[edge];
      }
      if (mode == 1) {
        mixture(variant, nc, d.class_logprob, use, logl, alpha);
        double m1 = 0.0, m2 = 0.0;
        for (int c = 0; c < nc; ++c) {
          m1 += alpha[c] * l1[c];
          m2 += alpha[c] * (l2[c] + l1[c] * l1[c]);
        }
        pd1 = m1;
        pd2 = m2 - m1 * m1;
      } else {
        int cnt = 0, only = -1;
        for (int c = 0; c < nc; ++c)
          if (use[c] && has_edge[c]) {
            ++cnt;
            only = c;
          }
        if (cnt == 1) {
          pd1 = l1[only];
          pd2 = l2[only];
        } else if (variant == 0) {
          double mx = -INFINITY;
          for (int c = 0; c < nc; ++c)
            if (use[c] && has_edge[c] && d.class_logprob[c] + logl[c] > mx)
              mx = d.class_logprob[c] + logl[c];
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
          for (int c = 0; c < nc; ++c) {
            if (!(use[c] && has_edge[c])) continue;
            double w = exp(d.class_logprob[c] + logl[c] - mx);
            s0 += w;
            s1 += w * l1[c];
            s2 += w * (l1[c] * l1[c] + l2[c]);
          }
          pd1 = s1 / s0;
          pd2 = (s2 * s0 - s1 * s1) / (s0 * s0);
        } else {
          double best = -INFINITY;
          for (int c = 0; c < nc; ++c) {
            if (!(use[c] && has_edge[c])) continue;
            double score = logl[c] * exp(d.class_logprob[c]);
            if (score > best) {
              best = score;
              pd1 = l1[c];
              pd2 = l2[c];
            }
          }
        }
      }
    }
    if (out_part_d1) out_part_d1[p] = pd1;
    if (out_part_d2) out_part_d2[p] = pd2;
    D1 += pd1;
    D2 += pd2;
    part_free(&ps, d.n_entries);
  }
  free(use);
  free(logl);
  free(l1);
  free(l2);
  free(alpha);
  free(t1);
  free(t2);
  free(has_edge);
  dag_free(&d);
  if (rc) return rc;
  *out_d1 = D1;
  *out_d2 = D2;
  return ORC_OK;
}
