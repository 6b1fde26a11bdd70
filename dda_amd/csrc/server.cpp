// Scattering-job server over the C ABI: a request queue, admission by HBM
// footprint, worker slots that each own a context (one HIP stream) so that
// admitted jobs run concurrently on the device, a shared device cache of FFT'd
// interaction tensors keyed by (grid, kd), and packing of queued jobs of the
// same particle (optionally: of any particle on the same grid) into one block
// solve. No counterpart in the reference, which
// runs one geometry per process (DDA/DDAMain.cpp:550-663).
//
// Continuous batching: a slot that finishes takes the next admissible job at
// once, while the other slots keep running, so requests that arrive while
// others are on the GPU start as soon as a slot and memory are free.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dda_hip.h"

namespace dda {
void set_error(const std::string& msg);  // api.hip: the message dda_last_error returns
bool last_tensor_cache_hit();            // matvec.hip: this thread's last cached-tensor lookup hit
double tensor_cache_seconds();           // matvec.hip: this thread's seconds in tensor-cache lookups
}

namespace {

struct Job {
  uint64_t id = 0;
  // geometry (owned copies)
  uint32_t box[3] = {0, 0, 0}, grid[3] = {0, 0, 0};
  uint32_t nv = 0, stride = 0, nmat = 0;
  double grid_unit = 0, wave_num = 0;
  std::vector<uint32_t> pos;
  std::vector<uint8_t> mat;
  uint64_t key = 0;  // hash of the geometry and wave number (packing)
  int nrhs = 0;
  std::vector<double> cc, einc;
  // results
  std::vector<double> pol;
  std::vector<dda_solve_info> info;
  dda_job_result res{};
  std::string error;  // message of a failed call in the worker
  bool done = false;
  bool claimed = false;  // a dda_server_wait call holds this job
};

uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; i++) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

// Jobs that can share a pack operator (dda_op_create_pack): one grid and
// grid unit, and room for the void material.
bool same_grid(const Job& a, const Job& b) {
  return a.grid_unit == b.grid_unit && std::equal(a.grid, a.grid + 3, b.grid) && a.nmat < 256 && b.nmat < 256;
}

bool same_particle(const Job& a, const Job& b) {
  return a.key == b.key && a.nv == b.nv && a.stride == b.stride && a.nmat == b.nmat &&
         a.grid_unit == b.grid_unit && a.wave_num == b.wave_num && std::equal(a.box, a.box + 3, b.box) &&
         std::equal(a.grid, a.grid + 3, b.grid) && a.pos == b.pos && a.mat == b.mat;
}

}  // namespace

struct dda_server_s {
  dda_server_config cfg{};
  dda_ctx_t ctx0 = nullptr;  // owns the tensor cache's device
  dda_dhat_cache_t cache = nullptr;
  size_t job_budget = 0;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_id = 1, submitted = 0, completed = 0;
  size_t in_use = 0, peak = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  // Device bytes a block solve of k right-hand sides of this geometry holds
  // (grid buffers of the operator and its FP32 twin, Krylov vectors, FP64
  // refinement vectors, index arrays); the tensor is counted by the cache.
  size_t footprint(const Job& j, int k) const {
    const size_t n = 3ull * j.stride;
    const size_t a = 3ull * j.box[2] * j.box[1] * j.grid[0] + 3ull * j.box[2] * j.grid[1] * j.grid[0];
    const bool lanczos = cfg.solver_kind == DDA_SOLVER_QMR_CS || cfg.solver_kind == DDA_SOLVER_BICG_CS;
    size_t b = 13ull * j.nv + 4ull * j.box[1] * j.box[2] + (size_t)k * a * 16;
    if (cfg.mode == DDA_MODE_MIXED) {
      b += (size_t)k * (a * 8 + 8 * n * 8 + 6 * n * 16);
      if (lanczos) b += (size_t)k * 8 * n * 16;  // FP64 inner fallback
    } else {
      b += (size_t)k * (8 * n * 16 + 4 * n * 16);
    }
    return b;
  }

  // Footprint of a pack of distinct particles: the union box and at most
  // the sum of their dipoles.
  size_t footprint(const std::vector<const Job*>& parts, int k) const {
    if (parts.size() == 1) return footprint(*parts[0], k);
    Job u;
    for (const Job* p : parts) {
      for (int a = 0; a < 3; a++) {
        u.box[a] = std::max(u.box[a], p->box[a]);
        u.grid[a] = p->grid[a];
      }
      u.nv += p->nv;
    }
    u.stride = (u.nv + 15) / 16 * 16;
    return footprint(u, k) + parts.size() * 5 * u.nv;
  }

  void run_pack(dda_ctx_t ctx, std::vector<Job*>& pack);
  void run_multi(dda_ctx_t ctx, std::vector<Job*>& pack, const std::vector<const Job*>& parts);
  void worker(int slot);
};

// Distinct particles of one grid: a pack operator over the union of their
// positions; each job's fields scattered into the union layout and its
// polarisations gathered back.
void dda_server_s::run_multi(dda_ctx_t ctx, std::vector<Job*>& pack, const std::vector<const Job*>& parts) {
  const int nm = (int)parts.size();
  std::vector<dda_op_desc> descs(nm);
  for (int m = 0; m < nm; m++) {
    const Job& p = *parts[m];
    dda_op_desc& d = descs[m];
    std::memset(&d, 0, sizeof d);
    std::copy(p.box, p.box + 3, d.box);
    std::copy(p.grid, p.grid + 3, d.grid);
    d.nv = p.nv;
    d.pos = p.pos.data();
    d.mat = p.mat.data();
    d.nmat = p.nmat;
    d.grid_unit = p.grid_unit;
    d.wave_num = p.wave_num;
  }
  int k = 0;
  for (Job* j : pack) k += j->nrhs;
  dda_op_t op = nullptr;
  dda_solver_t s = nullptr;
  const double t0 = now(), tc0 = dda::tensor_cache_seconds();
  dda_status rc = dda_op_create_pack(ctx, descs.data(), nm, 64, cache, &op);
  const bool hit = rc == DDA_OK && dda::last_tensor_cache_hit();
  int nmember = 0;
  uint32_t nvu = 0, nmu = 0, S = 0;
  if (rc == DDA_OK) rc = dda_op_pack_info(op, &nmember, &nvu, &nmu);
  if (rc == DDA_OK) rc = dda_op_info(op, nullptr, &S, nullptr);
  std::vector<std::vector<uint32_t>> maps(nm);
  for (int m = 0; m < nm && rc == DDA_OK; m++) {
    maps[m].resize(parts[m]->nv);
    rc = dda_op_pack_map(op, m, maps[m].data());
  }
  const size_t n = 3ull * S;
  std::vector<double> cc(2ull * 3 * nmu * k, 0.0), einc(2 * n * k, 0.0), pol(2 * n * k);
  std::vector<int> member(k);
  std::vector<dda_solve_info> info(k);
  std::vector<int> job_member(pack.size());
  if (rc == DDA_OK) {
    int r0 = 0;
    for (size_t q = 0; q < pack.size(); q++) {
      const Job& j = *pack[q];
      int m = 0;
      while (!same_particle(j, *parts[m])) m++;
      job_member[q] = m;
      for (int r = 0; r < j.nrhs; r++, r0++) {
        member[r0] = m;
        std::copy(j.cc.begin() + 6ull * j.nmat * r, j.cc.begin() + 6ull * j.nmat * (r + 1),
                  cc.begin() + 6ull * nmu * r0);
        for (int c = 0; c < 3; c++)
          for (uint32_t i = 0; i < j.nv; i++) {
            const size_t src = 2 * ((size_t)r * 3 * j.stride + (size_t)c * j.stride + i);
            const size_t dst = 2 * ((size_t)r0 * n + (size_t)c * S + maps[m][i]);
            einc[dst] = j.einc[src];
            einc[dst + 1] = j.einc[src + 1];
          }
      }
    }
  }
  if (rc == DDA_OK) rc = dda_op_set_cc_sqrt_batch(op, cc.data(), k);
  if (rc == DDA_OK) rc = dda_op_set_rhs_members(op, member.data(), k);
  if (rc == DDA_OK) rc = dda_solver_create_ex(op, cfg.solver_kind, cfg.max_iter, 0, cfg.mode, &s);
  if (rc == DDA_OK && cfg.long_solve_share > 0 && cfg.long_solve_share < 1)
    rc = dda_solver_set_gpu_share(s, cfg.long_solve_s, cfg.long_solve_share);
  const double setup = now() - t0;
  const double tensor_s = dda::tensor_cache_seconds() - tc0;  // FP64 tensor and (mixed) its FP32 twin
  if (rc == DDA_OK) rc = dda_solver_solve_batch(s, einc.data(), k, cfg.eps, nullptr, pol.data(), info.data());
  const bool call_failed = rc == DDA_HIP_ERROR || rc == DDA_OUT_OF_MEMORY || rc == DDA_NOT_SUPPORTED ||
                           (rc == DDA_INVALID_ARG && !s);
  const std::string err = rc == DDA_OK ? std::string() : std::string(dda_last_error());
  if (s) dda_solver_destroy(s);
  if (op) dda_op_destroy(op);
  int r0 = 0;
  for (size_t q = 0; q < pack.size(); q++) {
    Job* j = pack[q];
    const int m = job_member[q];
    j->pol.assign(6ull * j->stride * j->nrhs, 0.0);
    if (!call_failed)
      for (int r = 0; r < j->nrhs; r++)
        for (int c = 0; c < 3; c++)
          for (uint32_t i = 0; i < j->nv; i++) {
            const size_t src = 2 * ((size_t)(r0 + r) * n + (size_t)c * S + maps[m][i]);
            const size_t dst = 2 * ((size_t)r * 3 * j->stride + (size_t)c * j->stride + i);
            j->pol[dst] = pol[src];
            j->pol[dst + 1] = pol[src + 1];
          }
    j->info.assign(info.begin() + r0, info.begin() + r0 + j->nrhs);
    int st = DDA_OK;
    for (int r = 0; r < j->nrhs; r++)
      if (info[r0 + r].status != DDA_OK && st == DDA_OK) st = info[r0 + r].status;
    j->res.status = call_failed ? (int)rc : st;
    j->res.packed_rhs = k;
    j->res.cache_hit = hit ? 1 : 0;
    j->res.setup_s = setup;
    j->res.tensor_s = tensor_s;
    j->res.iterations = 0;
    for (int r = 0; r < j->nrhs; r++) j->res.iterations = std::max<int>(j->res.iterations, (int)info[r0 + r].iterations);
    j->error = err;
    r0 += j->nrhs;
  }
}

void dda_server_s::run_pack(dda_ctx_t ctx, std::vector<Job*>& pack) {
  std::vector<const Job*> parts;
  for (Job* j : pack)
    if (std::none_of(parts.begin(), parts.end(), [&](const Job* p) { return same_particle(*p, *j); }))
      parts.push_back(j);
  if (parts.size() > 1) {
    run_multi(ctx, pack, parts);
    return;
  }
  const Job& j0 = *pack[0];
  int k = 0;
  for (Job* j : pack) k += j->nrhs;
  const size_t n = 3ull * j0.stride;
  std::vector<double> cc, einc, pol(2 * n * k);
  std::vector<dda_solve_info> info(k);
  for (Job* j : pack) {
    cc.insert(cc.end(), j->cc.begin(), j->cc.end());
    einc.insert(einc.end(), j->einc.begin(), j->einc.end());
  }
  dda_op_desc d;
  std::memset(&d, 0, sizeof d);
  std::copy(j0.box, j0.box + 3, d.box);
  std::copy(j0.grid, j0.grid + 3, d.grid);
  d.nv = j0.nv;
  d.vec_stride = j0.stride;
  d.pos = j0.pos.data();
  d.mat = j0.mat.data();
  d.nmat = j0.nmat;
  d.grid_unit = j0.grid_unit;
  d.wave_num = j0.wave_num;
  dda_op_t op = nullptr;
  dda_solver_t s = nullptr;
  const double t0 = now(), tc0 = dda::tensor_cache_seconds();
  dda_status rc = dda_op_create_cached(ctx, &d, 64, cache, &op);
  const bool hit = rc == DDA_OK && dda::last_tensor_cache_hit();
  if (rc == DDA_OK) rc = dda_op_set_cc_sqrt_batch(op, cc.data(), k);
  if (rc == DDA_OK) rc = dda_solver_create_ex(op, cfg.solver_kind, cfg.max_iter, 0, cfg.mode, &s);
  if (rc == DDA_OK && cfg.long_solve_share > 0 && cfg.long_solve_share < 1)
    rc = dda_solver_set_gpu_share(s, cfg.long_solve_s, cfg.long_solve_share);
  const double setup = now() - t0;
  const double tensor_s = dda::tensor_cache_seconds() - tc0;  // FP64 tensor and (mixed) its FP32 twin
  if (rc == DDA_OK) rc = dda_solver_solve_batch(s, einc.data(), k, cfg.eps, nullptr, pol.data(), info.data());
  const bool call_failed = rc == DDA_HIP_ERROR || rc == DDA_OUT_OF_MEMORY || rc == DDA_NOT_SUPPORTED ||
                           (rc == DDA_INVALID_ARG && !s);
  const std::string err = rc == DDA_OK ? std::string() : std::string(dda_last_error());
  if (s) dda_solver_destroy(s);
  if (op) dda_op_destroy(op);
  int r0 = 0;
  for (Job* j : pack) {
    j->pol.assign(pol.begin() + 2 * n * r0, pol.begin() + 2 * n * (r0 + j->nrhs));
    j->info.assign(info.begin() + r0, info.begin() + r0 + j->nrhs);
    int st = DDA_OK;
    for (int r = 0; r < j->nrhs; r++)
      if (info[r0 + r].status != DDA_OK && st == DDA_OK) st = info[r0 + r].status;
    j->res.status = call_failed ? (int)rc : st;
    j->res.packed_rhs = k;
    j->res.cache_hit = hit ? 1 : 0;
    j->res.setup_s = setup;
    j->res.tensor_s = tensor_s;
    j->res.iterations = 0;
    for (int r = 0; r < j->nrhs; r++) j->res.iterations = std::max<int>(j->res.iterations, (int)info[r0 + r].iterations);
    j->error = err;
    r0 += j->nrhs;
  }
}

void dda_server_s::worker(int slot) {
  (void)slot;
  dda_ctx_t ctx = nullptr;  // each slot: its own stream, its device bound on this thread
  if (dda_ctx_create(cfg.device, &ctx) != DDA_OK) ctx = nullptr;
  for (;;) {
    std::vector<Job*> pack;
    size_t foot = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (stop && queue.empty()) break;
        if (!queue.empty()) {
          Job* h = queue.front();
          const size_t f = footprint(*h, h->nrhs);
          if (in_use == 0 || in_use + f <= job_budget) break;
        }
        cv_work.wait(lk);
      }
      if (queue.empty()) break;
      // the head job and queued jobs of the same particle (with pack_grids,
      // of any particle on its grid), up to max_pack RHS
      Job* h = queue.front();
      queue.pop_front();
      pack.push_back(h);
      int k = h->nrhs;
      std::vector<const Job*> parts{h};
      const size_t cap = std::max(job_budget, footprint(*h, h->nrhs));
      for (auto it = queue.begin(); it != queue.end() && k < cfg.max_pack;) {
        const Job& j = **it;
        bool fits = k + j.nrhs <= cfg.max_pack;
        bool is_new = false;
        if (fits) {
          is_new = std::none_of(parts.begin(), parts.end(), [&](const Job* p) { return same_particle(*p, j); });
          fits = !is_new || (cfg.pack_grids && same_grid(j, *h) &&
                             std::all_of(parts.begin(), parts.end(), [&](const Job* p) { return p->nmat < 256; }));
        }
        if (fits) {
          std::vector<const Job*> np = parts;
          if (is_new) np.push_back(&j);
          fits = in_use + footprint(np, k + j.nrhs) <= cap;
          if (fits) parts.swap(np);
        }
        if (fits) {
          k += j.nrhs;
          pack.push_back(*it);
          it = queue.erase(it);
        } else {
          ++it;
        }
      }
      foot = footprint(parts, k);
      in_use += foot;
      peak = std::max(peak, in_use);
      const double t = now();
      for (Job* j : pack) {
        j->res.t_start = t;
        j->res.footprint = foot;
      }
    }
    if (ctx) {
      run_pack(ctx, pack);
    } else {
      for (Job* j : pack) j->res.status = DDA_HIP_ERROR;
    }
    {
      std::lock_guard<std::mutex> g(mu);
      in_use -= foot;
      const double t = now();
      for (Job* j : pack) {
        j->res.t_end = t;
        j->done = true;
        completed++;
      }
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
  if (ctx) dda_ctx_destroy(ctx);
}

extern "C" dda_status dda_server_create(const dda_server_config* cfg, dda_server_t* out) {
  if (!cfg || !out || cfg->slots < 1 || cfg->slots > 16 || cfg->max_pack < 1 || !(cfg->eps > 0) ||
      cfg->solver_kind < 0 || cfg->solver_kind > 3 || (cfg->mode != DDA_MODE_FP64 && cfg->mode != DDA_MODE_MIXED) ||
      !(cfg->long_solve_s >= 0) || !(cfg->long_solve_share >= 0 && cfg->long_solve_share <= 1)) {
    dda::set_error("dda_server_create: need 1..16 slots, max_pack >= 1, eps > 0, a solver kind and mode");
    return DDA_INVALID_ARG;
  }
  *out = nullptr;
  dda_server_s* s = new dda_server_s();
  s->cfg = *cfg;
  s->t0 = std::chrono::steady_clock::now();
  dda_status rc = dda_ctx_create(cfg->device, &s->ctx0);
  size_t hbm = 0;
  if (rc == DDA_OK) rc = dda_ctx_info(s->ctx0, nullptr, 0, &hbm, nullptr);
  if (rc != DDA_OK) {
    if (s->ctx0) dda_ctx_destroy(s->ctx0);
    delete s;
    return rc;
  }
  const size_t budget = cfg->hbm_budget ? cfg->hbm_budget : hbm / 10 * 9;
  const size_t cache_bytes = cfg->cache_bytes ? cfg->cache_bytes : budget / 4;
  s->job_budget = budget > cache_bytes ? budget - cache_bytes : budget / 2;
  rc = dda_dhat_cache_create(s->ctx0, cache_bytes, &s->cache);
  if (rc != DDA_OK) {
    dda_ctx_destroy(s->ctx0);
    delete s;
    return rc;
  }
  for (int i = 0; i < cfg->slots; i++) s->workers.emplace_back(&dda_server_s::worker, s, i);
  *out = s;
  return DDA_OK;
}

extern "C" dda_status dda_server_submit(dda_server_t s, const dda_job_desc* job, uint64_t* id) {
  if (!s || !job || !id || job->nrhs < 1 || !job->cc_sqrt || !job->einc || job->geometry.dmatrix) {
    dda::set_error("dda_server_submit: null argument, nrhs < 1, or a host tensor was given");
    return DDA_INVALID_ARG;
  }
  const dda_op_desc& g = job->geometry;
  if (g.nv == 0 || !g.pos || g.nmat == 0 || g.nmat > 256 || !(g.grid_unit > 0) || !(g.wave_num > 0)) {
    dda::set_error("dda_server_submit: need nv > 0, positions, 1..256 materials, grid_unit > 0, wave_num > 0");
    return DDA_INVALID_ARG;
  }
  std::unique_ptr<Job> j(new Job());
  for (int a = 0; a < 3; a++) {
    j->box[a] = g.box[a];
    j->grid[a] = g.grid[a] ? g.grid[a] : dda_fft_fit_static(2 * g.box[a]);
    if (g.box[a] == 0 || j->grid[a] < 2 * g.box[a] || j->grid[a] % 2) {
      dda::set_error("dda_server_submit: grid must be even and at least twice the box");
      return DDA_INVALID_ARG;
    }
  }
  j->nv = g.nv;
  j->stride = g.vec_stride ? g.vec_stride : (g.nv + 15) / 16 * 16;
  if (j->stride < g.nv) {
    dda::set_error("dda_server_submit: vec_stride < nv");
    return DDA_INVALID_ARG;
  }
  j->nmat = g.nmat;
  j->grid_unit = g.grid_unit;
  j->wave_num = g.wave_num;
  j->pos.assign(g.pos, g.pos + 3ull * g.nv);
  j->mat.assign(g.nv, 0);
  if (g.mat) j->mat.assign(g.mat, g.mat + g.nv);
  j->nrhs = job->nrhs;
  j->cc.assign(job->cc_sqrt, job->cc_sqrt + 6ull * g.nmat * job->nrhs);
  j->einc.assign(job->einc, job->einc + 6ull * j->stride * job->nrhs);
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, j->box, sizeof j->box);
  h = fnv(h, j->grid, sizeof j->grid);
  h = fnv(h, &j->wave_num, sizeof j->wave_num);
  h = fnv(h, &j->grid_unit, sizeof j->grid_unit);
  h = fnv(h, j->pos.data(), j->pos.size() * 4);
  h = fnv(h, j->mat.data(), j->mat.size());
  j->key = h;
  {
    std::lock_guard<std::mutex> g2(s->mu);
    if (s->stop) {
      dda::set_error("dda_server_submit: the server is shutting down");
      return DDA_INVALID_ARG;
    }
    j->id = s->next_id++;
    j->res.t_submit = s->now();
    *id = j->id;
    s->queue.push_back(j.get());
    s->jobs[j->id] = std::move(j);
    s->submitted++;
  }
  s->cv_work.notify_all();
  return DDA_OK;
}

extern "C" dda_status dda_server_wait(dda_server_t s, uint64_t id, double* pol_out, dda_solve_info* info,
                                      dda_job_result* res) {
  if (!s) {
    dda::set_error("dda_server_wait: null server");
    return DDA_INVALID_ARG;
  }
  std::unique_ptr<Job> j;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) {
      dda::set_error("dda_server_wait: unknown job (or already collected)");
      return DDA_INVALID_ARG;
    }
    Job* p = it->second.get();
    if (p->claimed) {  // the first waiter frees the job when it returns
      dda::set_error("dda_server_wait: another thread is already waiting for this job");
      return DDA_INVALID_ARG;
    }
    p->claimed = true;
    s->cv_done.wait(lk, [&] { return p->done; });
    j = std::move(it->second);
    s->jobs.erase(it);
  }
  if (pol_out) std::copy(j->pol.begin(), j->pol.end(), pol_out);
  if (info) std::copy(j->info.begin(), j->info.end(), info);
  if (res) *res = j->res;
  if (j->res.status != DDA_OK) dda::set_error(j->error.empty() ? "job failed" : j->error);
  return (dda_status)j->res.status;
}

extern "C" dda_status dda_server_stats(dda_server_t s, uint64_t counts[5], size_t* peak_bytes) {
  if (!s) {
    dda::set_error("dda_server_stats: null server");
    return DDA_INVALID_ARG;
  }
  uint64_t c[3] = {0, 0, 0};
  dda_dhat_cache_stats(s->cache, c, nullptr, nullptr);
  std::lock_guard<std::mutex> g(s->mu);
  if (counts) {
    counts[0] = s->submitted;
    counts[1] = s->completed;
    counts[2] = c[0];
    counts[3] = c[1];
    counts[4] = c[2];
  }
  if (peak_bytes) *peak_bytes = s->peak;
  return DDA_OK;
}

extern "C" dda_status dda_server_destroy(dda_server_t s) {
  if (!s) return DDA_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto& t : s->workers) t.join();
  dda_dhat_cache_destroy(s->cache);
  dda_ctx_destroy(s->ctx0);
  delete s;
  return DDA_OK;
}
