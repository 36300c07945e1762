// Batched fixed-point solves: many right-hand sides (load cases of one
// solver, solvers of different grids) run to convergence in one sequence of
// launches.
//
// Entries on the same grid, Green operator kind and h, with the same
// schedule and rhs use, form a group.  A group is cut into chunks of `width`
// entries; every pass of a chunk is one launch of the batched register
// kernels (the block index selects the entry's slot), so small grids fill
// the GPU.  Chunks are cut from the running entries only: the device table
// of slots holds the running entries group after group and is rebuilt when
// entries stop, so that converged entries do not keep launching blocks.  A
// replay runs `iters_per_launch` iterations of every chunk, chunk after
// chunk, so that a chunk's working set stays in cache across its passes and
// iterations.  The first replay runs eagerly; with autotuning the next
// replays time the candidate widths of each group (hipEvents around the
// group's launches) and keep the fastest; the others are captured in a
// hipGraph and replayed, and the table and graph are rebuilt once a tenth
// of the running entries have stopped.  The host reads the solve states
// once per replay.  Each entry runs exactly the kernels and
// arithmetic of jx_solver_solve, and its converged blocks skip their work,
// so the iterates and iteration counts are those of the sequential solves.
//
// Conjugate-gradient entries (jx_batch_add_cg) run jx_solver_solve_cg: each
// starts and finishes on its own (the fill, the first matvec and residual;
// the final eps = E + x), and every iteration of a chunk is the fused CG
// iteration of the single solve with the chunk's entries in each launch,
// their fields, scalars and C0 weights in their slots.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "jx_handles.h"
#include "jx_internal.h"

using namespace jx;

namespace {

struct Entry {
  jx_solver *S = nullptr;
  double *eps = nullptr;
  const double *rhs = nullptr;
  double E[6] = {};
  double tol = 0, res_scale = 0;
  int max_iter = 0;
  double *spec = nullptr, *partials = nullptr, *res_hist = nullptr;
  int slot = -1;  // index of its solve state in the last run
  int iters = 0, status = 0;
  // conjugate gradients: r, p, q fields, vector partial sums, scalars and
  // the C0 inner product weights
  bool cg = false;
  double *cg_buf = nullptr, *cg_partials = nullptr;
  CgScalars *cg_cs = nullptr;
  double wd = 0., wt = 0.;
};

struct Group {
  std::vector<int> members;  // entry indices, in order of addition
  int off = 0;               // first solve state
  std::vector<int> act;      // entries in the device table (running)
  int act_off = 0;           // their first position in the table
  bool batched = false;      // register kernels take the whole chunk
  bool fused = false;
  bool cg = false;           // conjugate-gradient entries
  int width = 0;             // entries per launch (0 = all)
  bool user_width = false;
  // autotuning of the last run: candidate widths, ms per entry-iteration
  std::vector<int> tried;
  std::vector<double> cost;
};

}  // namespace

struct jx_batch {
  std::vector<Entry> ents;
  std::vector<Group> groups;
  std::map<int, int> width_set;  // group -> width chosen by the caller
  int iters_per_launch = 8;
  bool graph = true;
  bool autotune = false;
  // device state of the last run
  std::vector<Slot> slots;    // host copy of every entry's slot
  int built = 0;              // entries in the device table
  Slot *tab = nullptr;
  SolveState *st = nullptr;
  SolveState *st_host = nullptr;
  int cap = 0;
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  // concurrent groups: the groups' launches of a replay go to `nstreams`
  // streams (side streams forked from and joined back into `stream`)
  int nstreams = 1;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> side_done;
  hipEvent_t fork = nullptr;
  std::vector<int> group_stream;  // stream of each group, < nstreams
  // statistics of the last run
  double ms = 0;
  int replays = 0, captures = 0;
};

namespace {

using Key = std::tuple<int, int64_t, int64_t, int64_t, int, uint64_t, bool,
                       bool, int, int, bool>;

// Entries share launches when they have the same grid, kind, h (the axis
// tables), rhs use, schedule, plan variants (the rounding), family of
// local operator (the kernel instantiation) and method.
Key entry_key(const Entry &e) {
  const jx_green *G = e.S->green;
  uint64_t hb = 0;
  std::memcpy(&hb, &G->h, sizeof(hb));
  bool fused = e.S->fused && rows_fusable(G->g, G->sym);
  int plans = 0;
  for (int p = 0; p < kPlanPasses; ++p) plans |= G->g.plan[p] << p;
  return Key(G->g.dim, G->g.n[0], G->g.n[1], G->g.n[2], G->dev.kind, hb,
             e.rhs != nullptr, fused && !e.cg, plans,
             local_family(e.S->local->dev), e.cg);
}

void build_groups(jx_batch *B) {
  std::map<Key, int> index;
  B->groups.clear();
  for (int i = 0; i < (int)B->ents.size(); ++i) {
    Key k = entry_key(B->ents[i]);
    auto it = index.find(k);
    if (it == index.end()) {
      it = index.emplace(k, (int)B->groups.size()).first;
      Group gr;
      const jx_green *G = B->ents[i].S->green;
      gr.batched = batchable(G->g, G->sym);
      gr.fused = std::get<7>(k);
      gr.cg = std::get<10>(k);
      B->groups.push_back(gr);
    }
    B->groups[it->second].members.push_back(i);
  }
  int off = 0;
  for (int g = 0; g < (int)B->groups.size(); ++g) {
    Group &gr = B->groups[g];
    gr.off = off;
    off += (int)gr.members.size();
    auto w = B->width_set.find(g);
    gr.user_width = w != B->width_set.end();
    gr.width = gr.user_width ? w->second : 0;
    if (!gr.batched) gr.width = 1;
  }
}

// Entries per launch over n running entries of a group.
int chunk_width(const Group &gr, int width, int n) {
  if (!gr.batched) return 1;
  return (width <= 0 || width > n) ? std::max(n, 1) : width;
}

// One iteration of the chunk of group gr holding its table entries
// c0 ... c0 + n - 1.
int enqueue_chunk_iteration(jx_batch *B, const Group &gr, int c0, int n,
                            bool first, hipStream_t s) {
  int rc;
  if (gr.cg) {
    const Entry &e0 = B->ents[gr.act[c0]];
    Batch b;
    b.tab = B->tab + gr.act_off + c0;
    b.n = n;
    return cg_iteration_batched(e0.S->green, e0.S->local->dev, b, s);
  }
  if (!gr.batched) {
    for (int m = c0; m < c0 + n; ++m) {
      Entry &e = B->ents[gr.act[m]];
      RowsInverseLs ls;
      ls.enabled = true;
      ls.rhs = e.rhs;
      ls.partials = e.partials;
      SolveState *st = B->st + e.slot;
      int nb = 0;
      if ((rc = solver_iteration(e.S, e.eps, e.spec, st, ls, first, &nb, s)))
        return rc;
      if ((rc = launch_finalize(e.partials, nb, st, e.res_hist, s))) return rc;
    }
    return JX_OK;
  }
  const Entry &e0 = B->ents[gr.act[c0]];
  const jx_green *G = e0.S->green;
  const Grid &g = G->g;
  const int C = G->sym;
  const LocalDev &loc = e0.S->local->dev;
  Batch b;
  b.tab = B->tab + gr.act_off + c0;
  b.n = n;
  RowsInverseLs ls;
  ls.enabled = true;
  ls.rhs = e0.rhs;  // selects the kernel variant; the slots hold the fields
  int nb = 0;
  if ((!gr.fused || first) &&
      (rc = launch_rows_forward(g, C, nullptr, nullptr, loc, nullptr, s, &b)))
    return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), -1, nullptr, nullptr, nullptr, nullptr, s, &b)))
    return rc;
  if ((rc = launch_axis(g, C, pivot_axis(g), 0, nullptr, nullptr, &G->dev, nullptr, s, &b)))
    return rc;
  if (g.dim == 3 &&
      (rc = launch_axis(g, C, plain_axis(g), +1, nullptr, nullptr, nullptr, nullptr, s, &b)))
    return rc;
  if (gr.fused)
    rc = launch_rows_update_forward(g, C, nullptr, nullptr, ls, loc, nullptr,
                                    &nb, s, &b);
  else
    rc = launch_rows_inverse(g, C, nullptr, nullptr, ls, nullptr, &nb, s, &b);
  if (rc) return rc;
  return launch_finalize(nullptr, nb, nullptr, nullptr, s, &b);
}

bool entry_running(const jx_batch *B, int i) {
  return B->st_host[B->ents[i].slot].done == 0;
}

int count_running(const jx_batch *B) {
  int n = 0;
  for (int i = 0; i < (int)B->ents.size(); ++i) n += entry_running(B, i);
  return n;
}

// iters iterations of every chunk of group gr's table entries that has a
// running entry.
int enqueue_group(jx_batch *B, const Group &gr, int width, int iters,
                  bool first, hipStream_t s) {
  const int n = (int)gr.act.size();
  const int w = chunk_width(gr, width, n);
  for (int c0 = 0; c0 < n; c0 += w) {
    int cn = std::min(w, n - c0);
    bool running = false;
    for (int m = c0; m < c0 + cn; ++m) running |= entry_running(B, gr.act[m]);
    if (!running) continue;
    for (int k = 0; k < iters; ++k) {
      int rc = enqueue_chunk_iteration(B, gr, c0, cn, first && k == 0, s);
      if (rc) return rc;
    }
  }
  return JX_OK;
}

// Side streams and events for nstreams concurrent group streams (default
// priority, non-blocking).
int ensure_side_streams(jx_batch *B) {
  const int ns = B->nstreams;
  if (!B->fork) JX_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
  while ((int)B->side.size() < ns) {
    hipStream_t st = nullptr;
    JX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    B->side.push_back(st);
    hipEvent_t ev = nullptr;
    JX_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    B->side_done.push_back(ev);
  }
  return JX_OK;
}

// Groups to streams, longest first onto the least loaded stream, the load
// of a group being its running entries times their voxels (the bytes of
// one iteration of each).
void assign_streams(jx_batch *B) {
  const int ng = (int)B->groups.size();
  const int ns = std::max(1, std::min(B->nstreams, ng));
  std::vector<std::pair<double, int>> order;
  for (int g = 0; g < ng; ++g) {
    const Group &gr = B->groups[g];
    double w = 0.;
    if (!gr.act.empty()) {
      const Grid &G = B->ents[gr.act[0]].S->green->g;
      w = (double)G.N * (double)gr.act.size();
    }
    order.push_back({-w, g});
  }
  std::sort(order.begin(), order.end());
  std::vector<double> load(ns, 0.);
  B->group_stream.assign(ng, 0);
  for (const auto &o : order) {
    int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    B->group_stream[o.second] = j;
    load[j] -= o.first;
  }
}

// One replay of every group: on s alone, or (concurrent) forked to the side
// streams (each group on its assigned stream, its chunks in order) and
// joined back into s.  Groups share no fields, so the order of their
// launches changes no result.
int enqueue_replay(jx_batch *B, int iters, hipStream_t s, bool concurrent) {
  int rc = JX_OK;
  const int ng = (int)B->groups.size();
  if (!concurrent || B->nstreams <= 1 || ng <= 1) {
    for (const Group &gr : B->groups)
      if ((rc = enqueue_group(B, gr, gr.width, iters, false, s))) return rc;
    return JX_OK;
  }
  const int ns = std::min(B->nstreams, ng);
  JX_HIP(hipEventRecord(B->fork, s));
  for (int j = 0; j < ns; ++j) JX_HIP(hipStreamWaitEvent(B->side[j], B->fork, 0));
  for (int g = 0; g < ng && !rc; ++g)
    rc = enqueue_group(B, B->groups[g], B->groups[g].width, iters, false,
                       B->side[B->group_stream[g]]);
  // join even after a failed enqueue, so that s waits for what was queued
  for (int j = 0; j < ns; ++j) {
    JX_HIP(hipEventRecord(B->side_done[j], B->side[j]));
    JX_HIP(hipStreamWaitEvent(s, B->side_done[j], 0));
  }
  return rc;
}

// Device table of the running entries, group after group.
int upload_table(jx_batch *B, hipStream_t s) {
  std::vector<Slot> tab;
  for (Group &gr : B->groups) {
    gr.act.clear();
    gr.act_off = (int)tab.size();
    for (int i : gr.members)
      if (entry_running(B, i)) {
        gr.act.push_back(i);
        tab.push_back(B->slots[i]);
      }
  }
  B->built = (int)tab.size();
  assign_streams(B);
  if (!tab.empty())
    JX_HIP(hipMemcpyAsync(B->tab, tab.data(), sizeof(Slot) * tab.size(),
                          hipMemcpyHostToDevice, s));
  JX_HIP(hipStreamSynchronize(s));
  return JX_OK;
}

int read_states(jx_batch *B, hipStream_t s) {
  const int n = (int)B->ents.size();
  JX_HIP(hipMemcpyAsync(B->st_host, B->st, sizeof(SolveState) * n,
                        hipMemcpyDeviceToHost, s));
  JX_HIP(hipStreamSynchronize(s));
  return JX_OK;
}

void drop_graph(jx_batch *B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  B->exec = nullptr;
}

// Captures one replay (all running chunks) into B->exec.
int capture(jx_batch *B, hipStream_t s) {
  drop_graph(B);
  JX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  // captured on s alone: a capture forked onto the side streams aborted the
  // process with conjugate-gradient entries, for a cause not found
  int rc = enqueue_replay(B, B->iters_per_launch, s, false);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "jx_batch_run: hipStreamEndCapture");
  e = hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    B->exec = nullptr;
    return hip_fail(e, "jx_batch_run: hipGraphInstantiate");
  }
  ++B->captures;
  return JX_OK;
}

std::vector<int> candidate_widths(const Group &gr) {
  const int n = (int)gr.members.size();
  std::vector<int> c;
  if (!gr.batched || gr.user_width || n <= 1) return c;
  for (int w : {n, 12, 6, 3, 2, 1})
    if (w <= n && std::find(c.begin(), c.end(), w) == c.end()) c.push_back(w);
  return c;
}

void free_entry(Entry &e) {
  if (e.spec) (void)hipFree(e.spec);
  if (e.partials) (void)hipFree(e.partials);
  if (e.res_hist) (void)hipFree(e.res_hist);
  if (e.cg_buf) (void)hipFree(e.cg_buf);
  if (e.cg_partials) (void)hipFree(e.cg_partials);
  if (e.cg_cs) (void)hipFree(e.cg_cs);
  e.spec = e.partials = e.res_hist = e.cg_buf = e.cg_partials = nullptr;
  e.cg_cs = nullptr;
}

// The work buffers of a conjugate-gradient entry (its solve state in the
// run's table).
CgWork entry_cg_work(const jx_batch *B, const Entry &e) {
  const Grid &g = e.S->green->g;
  const int64_t F = (int64_t)e.S->green->sym * g.N;
  CgWork w;
  w.r = e.cg_buf;
  w.p = e.cg_buf + F;
  w.q = e.cg_buf + 2 * F;
  w.spec = e.spec;
  w.rows_partials = e.partials;
  w.vec_partials = e.cg_partials;
  w.cs = e.cg_cs;
  w.st = B->st + e.slot;
  w.res_hist = e.res_hist;
  return w;
}

void free_run_state(jx_batch *B) {
  drop_graph(B);
  if (B->tab) (void)hipFree(B->tab);
  if (B->st) (void)hipFree(B->st);
  if (B->st_host) (void)hipHostFree(B->st_host);
  B->tab = nullptr;
  B->st = nullptr;
  B->st_host = nullptr;
  B->cap = 0;
}

int run(jx_batch *B, hipStream_t caller) {
  const int n = (int)B->ents.size();
  if (n == 0) return JX_OK;
  build_groups(B);
  if (B->cap < n) {
    free_run_state(B);
    JX_HIP(hipMalloc((void **)&B->tab, sizeof(Slot) * n));
    JX_HIP(hipMalloc((void **)&B->st, sizeof(SolveState) * n));
    JX_HIP(hipHostMalloc((void **)&B->st_host, sizeof(SolveState) * n, 0));
    B->cap = n;
  }
  drop_graph(B);
  if (!B->stream)
    JX_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
  if (B->nstreams > 1) {
    int rcs = ensure_side_streams(B);
    if (rcs) return rcs;
  }
  hipStream_t s = B->stream;
  JX_HIP(hipStreamSynchronize(caller));

  B->slots.assign(n, Slot());
  for (Group &gr : B->groups)
    for (int m = 0; m < (int)gr.members.size(); ++m) {
      Entry &e = B->ents[gr.members[m]];
      e.slot = gr.off + m;
      Slot &sl = B->slots[gr.members[m]];
      sl.eps = e.eps;
      sl.spec = e.spec;
      sl.rhs = e.rhs;
      sl.partials = e.partials;
      sl.res_hist = e.res_hist;
      sl.st = B->st + e.slot;
      sl.loc = e.S->local->dev;
      const GreenDev &D = e.S->green->dev;
      sl.daux[0] = D.daux1;
      sl.daux[1] = D.daux2;
      sl.daux[2] = D.daux3;
      sl.daux[3] = D.daux4;
      if (e.cg) {
        const CgWork w = entry_cg_work(B, e);
        sl.cg_r = w.r;
        sl.cg_p = w.p;
        sl.cg_q = w.q;
        sl.cg_cs = e.cg_cs;
        sl.cg_partials = e.cg_partials;
        sl.cg_wd = e.wd;
        sl.cg_wt = e.wt;
      }
      SolveState &st = B->st_host[e.slot];
      st.iter = 0;
      st.done = 0;
      st.tol = e.tol;
      st.res_scale = e.res_scale;
      st.max_iter = e.max_iter;
      st.pad = 0;
      for (int c = 0; c < 6; ++c) st.E[c] = e.E[c];
    }
  JX_HIP(hipMemcpyAsync(B->st, B->st_host, sizeof(SolveState) * n,
                        hipMemcpyHostToDevice, s));
  int rc = upload_table(B, s);
  if (rc) return rc;
  // conjugate-gradient entries start one by one (x = 0, r = p = rhs -
  // Gamma0[dC : E], the first residual, which stops a homogeneous medium)
  for (const Entry &e : B->ents)
    if (e.cg &&
        (rc = cg_start(e.S, entry_cg_work(B, e), e.eps, e.rhs, e.wd, e.wt, s)))
      return rc;
  B->replays = 0;
  B->captures = 0;
  auto t0 = std::chrono::steady_clock::now();
  const int K = B->iters_per_launch;

  // replay 0: eager, carries pass A of the fused groups
  for (Group &gr : B->groups) {
    std::vector<int> cand = candidate_widths(gr);
    if (!cand.empty()) gr.width = cand[0];
    if ((rc = enqueue_group(B, gr, gr.width, K, true, s))) return rc;
  }
  if ((rc = read_states(B, s))) return rc;
  ++B->replays;

  // autotuning replays: candidate r of every group, timed per group
  if (B->autotune) {
    const int ng = (int)B->groups.size();
    std::vector<std::vector<int>> cand(ng);
    std::vector<std::vector<double>> cost(ng);
    size_t rounds = 0;
    for (int g = 0; g < ng; ++g) {
      cand[g] = candidate_widths(B->groups[g]);
      rounds = std::max(rounds, cand[g].size());
    }
    std::vector<hipEvent_t> ev(2 * ng);
    for (auto &e : ev) JX_HIP(hipEventCreate(&e));
    for (size_t r = 0; r < rounds && !rc; ++r) {
      if (count_running(B) < B->built && (rc = upload_table(B, s))) break;
      // iterations done by the group's entries before the replay
      std::vector<long long> before(ng, 0);
      for (int g = 0; g < ng; ++g)
        for (int i : B->groups[g].members)
          before[g] += B->st_host[B->ents[i].slot].iter;
      for (int g = 0; g < ng && !rc; ++g) {
        Group &gr = B->groups[g];
        int w = r < cand[g].size() ? cand[g][r] : gr.width;
        (void)hipEventRecord(ev[2 * g], s);
        rc = enqueue_group(B, gr, w, K, false, s);
        (void)hipEventRecord(ev[2 * g + 1], s);
      }
      if (!rc) rc = read_states(B, s);
      ++B->replays;
      for (int g = 0; g < ng && !rc; ++g) {
        if (r >= cand[g].size()) continue;
        long long done = -before[g];
        for (int i : B->groups[g].members)
          done += B->st_host[B->ents[i].slot].iter;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[2 * g], ev[2 * g + 1]);
        // time per entry-iteration actually run (stopped entries' blocks
        // exit at once and count for nothing)
        cost[g].push_back(done > 0 ? ms / (double)done : 1e30);
      }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int g = 0; g < ng; ++g) {
      B->groups[g].tried = cand[g];
      B->groups[g].cost = cost[g];
      if (cost[g].empty()) continue;
      size_t best = 0;
      for (size_t i = 1; i < cost[g].size(); ++i)
        if (cost[g][i] < cost[g][best]) best = i;
      // no valid measurement: keep the whole group per launch
      B->groups[g].width = cost[g][best] < 1e30 ? cand[g][best] : cand[g][0];
    }
  }

  // steady state: replays until every entry has stopped; the table (and
  // the graph) is rebuilt once a tenth of its entries have stopped, or at
  // every change without a graph
  for (int running = count_running(B); running > 0;
       running = count_running(B)) {
    if (running < B->built &&
        (!B->graph || 10 * running <= 9 * B->built || !B->exec)) {
      drop_graph(B);
      if ((rc = upload_table(B, s))) return rc;
    }
    if (B->graph) {
      if (!B->exec && (rc = capture(B, s))) return rc;
      JX_HIP(hipGraphLaunch(B->exec, s));
    } else if ((rc = enqueue_replay(B, K, s, true))) {
      (void)hipStreamSynchronize(s);
      return rc;
    }
    if ((rc = read_states(B, s))) return rc;
    ++B->replays;
  }
  drop_graph(B);
  // eps = E + x + alpha p for the conjugate-gradient entries
  for (const Entry &e : B->ents)
    if (e.cg && (rc = cg_finish(e.S, entry_cg_work(B, e), e.eps, e.wd, e.wt, s)))
      return rc;
  JX_HIP(hipStreamSynchronize(s));
  B->ms = std::chrono::duration<double, std::milli>(
              std::chrono::steady_clock::now() - t0).count();
  for (Entry &e : B->ents) {
    const SolveState &st = B->st_host[e.slot];
    e.iters = st.iter;
    e.status = st.done ? st.done : JX_SOLVE_MAXITER;
  }
  return JX_OK;
}

}  // namespace

extern "C" {

int jx_batch_create(jx_batch **out) {
  if (!out) return fail(JX_EINVAL, "jx_batch_create: out is NULL");
  *out = new jx_batch();
  return JX_OK;
}

int jx_batch_destroy(jx_batch *B) {
  if (!B) return JX_OK;
  free_run_state(B);
  for (Entry &e : B->ents) free_entry(e);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  for (hipStream_t st : B->side) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : B->side_done) (void)hipEventDestroy(ev);
  if (B->fork) (void)hipEventDestroy(B->fork);
  delete B;
  return JX_OK;
}

static int batch_add(jx_batch *B, jx_solver *S, double *eps, const double *E,
                     const double *rhs, double tol, int max_iter,
                     double res_scale, int *index, bool cg) {
  if (!B || !S || !eps || !E)
    return fail(JX_EINVAL, "jx_batch_add: NULL argument");
  JX_ALIGNED(eps, "jx_batch_add");
  JX_ALIGNED(rhs, "jx_batch_add");
  if (max_iter < 1) return fail(JX_EINVAL, "max_iter must be >= 1");
  const jx_green *G = S->green;
  const int C = G->sym;
  if (cg) {
    if (!S->local->symmetric)
      return fail(JX_EINVAL, "conjugate gradients need a symmetric local operator");
    if (!cg_batchable(G->g, C))
      return fail(JX_EUNSUPPORTED, "batched conjugate gradients need register "
                                   "FFT lengths on every axis (use "
                                   "jx_solver_solve_cg)");
  }
  Entry e;
  e.S = S;
  e.cg = cg;
  if (cg) cg_weights(G, &e.wd, &e.wt);
  e.eps = eps;
  e.rhs = rhs;
  for (int c = 0; c < C; ++c) e.E[c] = E[c];
  e.tol = tol;
  e.max_iter = max_iter;
  if (res_scale <= 0.) {
    double nE = 0.;
    for (int c = 0; c < C; ++c) nE += E[c] * E[c];
    nE = std::sqrt(nE);
    if (nE == 0.) return fail(JX_EINVAL, "res_scale must be given when E = 0");
    res_scale = 1.0 / std::sqrt((double)G->g.N) / nE;  // square_basic.py:86
  }
  e.res_scale = res_scale;
  for (const Entry &o : B->ents)
    if (o.eps == eps)
      return fail(JX_EINVAL, "jx_batch_add: each entry needs its own eps");
  int nb = rows_inverse_nblocks(G->g, C);
  hipError_t he = hipMalloc((void **)&e.spec, sizeof(double) * spec_doubles(G->g, C));
  if (he == hipSuccess) he = hipMalloc((void **)&e.partials, sizeof(double) * nb);
  if (he == hipSuccess)
    he = hipMalloc((void **)&e.res_hist, sizeof(double) * max_iter);
  if (cg && he == hipSuccess)
    he = hipMalloc((void **)&e.cg_buf, sizeof(double) * 3 * C * G->g.N);
  if (cg && he == hipSuccess)
    he = hipMalloc((void **)&e.cg_partials,
                   sizeof(double) * cg_vec_partials_doubles());
  if (cg && he == hipSuccess)
    he = hipMalloc((void **)&e.cg_cs, sizeof(CgScalars));
  if (he != hipSuccess) {
    free_entry(e);
    return hip_fail(he, "jx_batch_add");
  }
  B->ents.push_back(e);
  if (index) *index = (int)B->ents.size() - 1;
  return JX_OK;
}

int jx_batch_add(jx_batch *B, jx_solver *S, double *eps, const double *E,
                 const double *rhs, double tol, int max_iter, double res_scale,
                 int *index) {
  return batch_add(B, S, eps, E, rhs, tol, max_iter, res_scale, index, false);
}

int jx_batch_add_cg(jx_batch *B, jx_solver *S, double *eps, const double *E,
                    const double *rhs, double tol, int max_iter,
                    double res_scale, int *index) {
  return batch_add(B, S, eps, E, rhs, tol, max_iter, res_scale, index, true);
}

int jx_batch_ngroups(jx_batch *B, int *ngroups) {
  if (!B || !ngroups) return fail(JX_EINVAL, "jx_batch_ngroups: NULL argument");
  std::vector<Group> keep = B->groups;
  build_groups(B);
  // keep the widths chosen by the last run
  if (keep.size() == B->groups.size())
    for (size_t g = 0; g < keep.size(); ++g)
      if (keep[g].members == B->groups[g].members) {
        B->groups[g].width = keep[g].width;
        B->groups[g].tried = keep[g].tried;
        B->groups[g].cost = keep[g].cost;
      }
  *ngroups = (int)B->groups.size();
  return JX_OK;
}

int jx_batch_group(jx_batch *B, int g, int *dim, int64_t *shape, int *size,
                   int *width, int *batched) {
  if (!B) return fail(JX_EINVAL, "jx_batch_group: NULL batch");
  if (g < 0 || g >= (int)B->groups.size())
    return fail(JX_EINVAL, "jx_batch_group: group out of range");
  const Group &gr = B->groups[g];
  const Grid &G = B->ents[gr.members[0]].S->green->g;
  if (dim) *dim = G.dim;
  if (shape)
    for (int d = 0; d < 3; ++d) shape[d] = d < G.dim ? G.n[d] : 1;
  if (size) *size = (int)gr.members.size();
  if (width) *width = chunk_width(gr, gr.width, (int)gr.members.size());
  if (batched) *batched = gr.batched ? 1 : 0;
  return JX_OK;
}

int jx_batch_tuning(jx_batch *B, int g, int cap, int *n, int *widths,
                    double *ms) {
  if (!B || !n) return fail(JX_EINVAL, "jx_batch_tuning: NULL argument");
  if (g < 0 || g >= (int)B->groups.size())
    return fail(JX_EINVAL, "jx_batch_tuning: group out of range");
  const Group &gr = B->groups[g];
  *n = (int)gr.cost.size();
  for (int i = 0; i < *n && i < cap; ++i) {
    if (widths) widths[i] = gr.tried[i];
    if (ms) ms[i] = gr.cost[i];
  }
  return JX_OK;
}

int jx_batch_set_width(jx_batch *B, int g, int width) {
  if (!B || g < 0) return fail(JX_EINVAL, "jx_batch_set_width: invalid argument");
  if (width < 0) return fail(JX_EINVAL, "width must be >= 0");
  B->width_set[g] = width;
  return JX_OK;
}

int jx_batch_set_schedule(jx_batch *B, int iters_per_launch, int graph,
                          int autotune) {
  if (!B) return fail(JX_EINVAL, "jx_batch_set_schedule: NULL batch");
  if (iters_per_launch < 1) return fail(JX_EINVAL, "iters_per_launch must be >= 1");
  B->iters_per_launch = iters_per_launch;
  B->graph = graph != 0;
  B->autotune = autotune != 0;
  return JX_OK;
}

int jx_batch_set_streams(jx_batch *B, int nstreams) {
  if (!B) return fail(JX_EINVAL, "jx_batch_set_streams: NULL batch");
  if (nstreams < 1 || nstreams > 8)
    return fail(JX_EINVAL, "nstreams must be in [1, 8]");
  B->nstreams = nstreams;
  return JX_OK;
}

int jx_batch_run(jx_batch *B, void *stream) {
  if (!B) return fail(JX_EINVAL, "jx_batch_run: NULL batch");
  int rc = run(B, (hipStream_t)stream);
  if (rc) {
    drop_graph(B);
    for (hipStream_t st : B->side) (void)hipStreamSynchronize(st);
    if (B->stream) (void)hipStreamSynchronize(B->stream);
  }
  return rc;
}

int jx_batch_result(jx_batch *B, int index, double *res_hist, int *iters,
                    int *status) {
  if (!B || index < 0 || index >= (int)B->ents.size())
    return fail(JX_EINVAL, "jx_batch_result: invalid entry");
  const Entry &e = B->ents[index];
  if (iters) *iters = e.iters;
  if (status) *status = e.status;
  if (res_hist && e.iters > 0)
    JX_HIP(hipMemcpy(res_hist, e.res_hist, sizeof(double) * e.iters,
                     hipMemcpyDeviceToHost));
  return JX_OK;
}

int jx_batch_stats(jx_batch *B, double *ms, int *replays, int *captures) {
  if (!B) return fail(JX_EINVAL, "jx_batch_stats: NULL batch");
  if (ms) *ms = B->ms;
  if (replays) *replays = B->replays;
  if (captures) *captures = B->captures;
  return JX_OK;
}

}  // extern "C"
