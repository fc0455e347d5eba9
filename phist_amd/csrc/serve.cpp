// serve.cpp -- continuous-batching service loop (include/phist_amd_serve.h).
//
// Per (matrix, precision) one MinresBatch with `slots` columns; queued
// shifted systems fill free slots in arrival order, every step advances each
// batch by one MINRES iteration (while eigenvalue requests run, by as many
// as fit in the time their outer iterations took in the last step, at most
// PHIST_AMD_SERVE_MAX_REPS) and harvests the systems that finished
// (their x is copied into the request's X column and the slot is released
// for the next system).  Eigenvalue requests hold a resumable block
// Jacobi-Davidson and advance one outer iteration per step.  Admission is
// first-come first-served against a byte budget for solver work space; a
// request whose work space can never fit fails with -44, and idle batches
// give their memory back when the queue head waits for it.  Queued or
// running requests can be cancelled.
#include "phist_amd_serve.h"
#include "solvers.hpp"
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <unordered_map>

namespace phx {
namespace {

double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

struct Request {
  int64_t id = 0;
  bool jada = false;
  int state = PHIST_AMD_REQ_QUEUED;
  double t_submit = 0, t_done = 0;
  int iters = 0;
  int64_t bytes = 0;  // work space charged while running
  // shifted solves
  const SparseMat<double>* A = nullptr;
  bool fp32 = false;
  std::vector<double> shifts;
  const Mvec<double>* B = nullptr;
  Mvec<double>* X = nullptr;
  double tol = 0;
  int max_iter = 0;
  int next_col = 0;   // first system not yet in a slot
  int remaining = 0;  // systems not yet finished
  bool all_converged = true;
  // eigenvalue requests
  phist_jadaOpts opts;
  Mvec<double>* Q = nullptr;
  SdMat<double>* R = nullptr;
  double* ev = nullptr;
  double* res = nullptr;
  JadaHandle* jd = nullptr;
};

// one MINRES batch over a matrix in one precision
struct Batch {
  const SparseMat<double>* A = nullptr;
  bool fp32 = false;
  Operator<double> op64;
  Operator<float> op32;
  std::unique_ptr<MinresBatch<double>> mb64;
  std::unique_ptr<MinresBatch<float>> mb32;
  std::vector<std::pair<Request*, int>> owner;  // per slot: request and its column
  int64_t bytes = 0;

  int free_slot() const { return fp32 ? mb32->free_slot() : mb64->free_slot(); }
  int n_used() const { return fp32 ? mb32->n_used() : mb64->n_used(); }
};

class Server {
 public:
  Server(int64_t budget, int slots) : budget_(budget), slots_(slots) {
    PHX_REQUIRE(slots >= 1 && slots <= 64, PHIST_INVALID_INPUT, "slots %d not in 1..64", slots);
    PHX_REQUIRE(budget > 0, PHIST_INVALID_INPUT, "budget must be positive");
  }
  ~Server() {
    for (auto& kv : reqs_)
      if (kv.second->jd) jada_free(kv.second->jd);
  }

  int64_t submit_solve(const SparseMat<double>* A, int nrhs, const double* shifts, const Mvec<double>* B,
                       Mvec<double>* X, double tol, int max_iter, bool fp32) {
    PHX_REQUIRE(A && B && X && nrhs >= 1 && B->nvec >= nrhs && X->nvec >= nrhs, PHIST_INVALID_INPUT,
                "bad solve request");
    PHX_REQUIRE(A->nrows == A->ncols && B->nrows == A->nrows && X->nrows == A->nrows, PHIST_INVALID_INPUT,
                "shapes do not match the square matrix");
    auto r = std::make_unique<Request>();
    r->id = ++last_id_;
    r->t_submit = now_s();
    r->A = A;
    r->fp32 = fp32;
    r->shifts.assign(shifts, shifts + nrhs);
    r->B = B;
    r->X = X;
    r->tol = tol;
    r->max_iter = max_iter;
    r->remaining = nrhs;
    Request* p = r.get();
    reqs_[p->id] = std::move(r);
    queue_.push_back(p);
    return p->id;
  }

  int64_t submit_jada(const SparseMat<double>* A, const phist_jadaOpts& opts, bool fp32, Mvec<double>* Q,
                      SdMat<double>* R, double* ev, double* res) {
    PHX_REQUIRE(A && Q && R && ev && res, PHIST_INVALID_INPUT, "bad eigenvalue request");
    // unsupported options fail at submission, as phist_Dsubspacejada fails
    jada_validate(opts, opts.symmetry == phist_GENERAL && sparse_is_symmetric(A));
    auto r = std::make_unique<Request>();
    r->id = ++last_id_;
    r->jada = true;
    r->t_submit = now_s();
    r->A = A;
    r->fp32 = fp32;
    r->opts = opts;
    r->Q = Q;
    r->R = R;
    r->ev = ev;
    r->res = res;
    const int64_t n = A->nrows, bs = opts.blockSize, nq = opts.numEigs + opts.blockSize - 1;
    r->bytes = (int64_t)8 * n * (2 * opts.maxBas + nq + bs + std::min<int64_t>(2 * bs, nq) + 8 * bs + nq + 4);
    Request* p = r.get();
    reqs_[p->id] = std::move(r);
    queue_.push_back(p);
    return p->id;
  }

  int step() {
    admit();
    // MINRES iterations per batch, then harvest finished systems: one, or, while eigenvalue requests
    // run, as many as took the time of their outer iterations in the last step (at most
    // serve_max_reps), so that
    // shifted solves do not advance at the pace of the eigensolver's steps
    int reps = 1;
    if (!running_jd_.empty() && t_jd_ > 0.0 && t_iter_ > 0.0)
      reps = (int)std::max(1.0, std::min((double)rt().serve_max_reps, std::floor(t_jd_ / t_iter_)));
    const double t0 = now_s();
    int iters = 0;
    for (auto& b : batches_) {
      for (int rep = 0; rep < reps && b->n_used() > 0; rep++) {
        if (b->fp32)
          b->mb32->iterate();
        else
          b->mb64->iterate();
        iters++;
        harvest(*b);
      }
    }
    if (iters > 0) t_iter_ = (now_s() - t0) / iters;  // per iteration of one batch, this step
    // one outer iteration per running eigenvalue request
    const double t1 = now_s();
    for (size_t i = 0; i < running_jd_.size();) {
      Request* r = running_jd_[i];
      bool fin = false;
      try {
        fin = jada_step(r->jd);
      } catch (const Error& e) {
        r->state = e.code < 0 ? e.code : PHIST_FUNCTIONAL_ERROR;
        fin = true;
      }
      if (fin) {
        if (r->state >= 0) {
          int nconv = 0, nit = 0;
          jada_finish(r->jd, &nconv, &nit);
          r->iters = nit;
          r->state = nconv >= r->opts.numEigs ? PHIST_AMD_REQ_DONE : PHIST_AMD_REQ_DONE_UNCONVERGED;
        }
        r->t_done = now_s();
        jada_free(r->jd);
        r->jd = nullptr;
        in_use_ -= r->bytes;
        running_jd_.erase(running_jd_.begin() + i);
      } else {
        i++;
      }
    }
    t_jd_ = running_jd_.empty() ? 0.0 : now_s() - t1;
    // drop idle batches (their work space returns to the budget)
    for (size_t i = 0; i < batches_.size();) {
      if (batches_[i]->n_used() == 0 && !queued_for(batches_[i]->A, batches_[i]->fp32)) {
        in_use_ -= batches_[i]->bytes;
        batches_.erase(batches_.begin() + i);
      } else {
        i++;
      }
    }
    int pending = 0;
    for (auto& kv : reqs_)
      if (kv.second->state == PHIST_AMD_REQ_QUEUED || kv.second->state == PHIST_AMD_REQ_RUNNING) pending++;
    return pending;
  }

  void status(int64_t id, int* state, double* latency, int* iters) const {
    auto it = reqs_.find(id);
    PHX_REQUIRE(it != reqs_.end(), PHIST_INVALID_INPUT, "unknown request %lld", (long long)id);
    const Request& r = *it->second;
    *state = r.state;
    *latency = r.state >= PHIST_AMD_REQ_DONE || r.state < 0 ? r.t_done - r.t_submit : now_s() - r.t_submit;
    *iters = r.iters;
  }

  // Withdraws a queued or running request: its systems leave their slots
  // (which the next step refills), a running eigenvalue request frees its
  // work space; its outputs are left as they are.  A finished request is
  // not changed.
  void cancel(int64_t id) {
    auto it = reqs_.find(id);
    PHX_REQUIRE(it != reqs_.end(), PHIST_INVALID_INPUT, "unknown request %lld", (long long)id);
    Request* r = it->second.get();
    if (r->state != PHIST_AMD_REQ_QUEUED && r->state != PHIST_AMD_REQ_RUNNING) return;
    for (auto q = queue_.begin(); q != queue_.end();)
      q = (*q == r) ? queue_.erase(q) : q + 1;
    if (r->jada) {
      if (r->jd) {
        jada_free(r->jd);
        r->jd = nullptr;
        in_use_ -= r->bytes;
        for (size_t i = 0; i < running_jd_.size(); i++)
          if (running_jd_[i] == r) {
            running_jd_.erase(running_jd_.begin() + i);
            break;
          }
      }
    } else {
      for (auto& b : batches_)
        for (int j = 0; j < slots_; j++)
          if (b->owner[j].first == r) {
            if (b->fp32)
              b->mb32->release(j);
            else
              b->mb64->release(j);
            b->owner[j] = {nullptr, 0};
          }
    }
    r->state = PHIST_AMD_REQ_CANCELLED;
    r->t_done = now_s();
  }

  void release(int64_t id) {
    auto it = reqs_.find(id);
    PHX_REQUIRE(it != reqs_.end(), PHIST_INVALID_INPUT, "unknown request %lld", (long long)id);
    const int st = it->second->state;
    PHX_REQUIRE(st != PHIST_AMD_REQ_QUEUED && st != PHIST_AMD_REQ_RUNNING, PHIST_INVALID_INPUT,
                "request %lld is not finished", (long long)id);
    reqs_.erase(it);
  }

 private:
  bool queued_for(const SparseMat<double>* A, bool fp32) const {
    for (const Request* r : queue_)
      if (!r->jada && r->A == A && r->fp32 == fp32) return true;
    return false;
  }

  Batch* find_batch(const SparseMat<double>* A, bool fp32) {
    for (auto& b : batches_)
      if (b->A == A && b->fp32 == fp32) return b.get();
    return nullptr;
  }

  // work space of a MINRES batch: its six blocks of slots columns (an FP32
  // copy of the matrix belongs to the matrix, i.e. to the caller's memory)
  int64_t batch_bytes(const SparseMat<double>* A, bool fp32) const {
    return (int64_t)(fp32 ? 4 : 8) * A->nrows * 6 * slots_;
  }

  // Drops batches with no system in flight until `need` more bytes fit the
  // budget.  Called for the queue head when it waits for memory: an idle
  // batch is kept at the end of a step while requests for it are queued,
  // but requests behind the head cannot use it before the head runs, so
  // holding it would block the head (and everything behind it) for good.
  void release_idle_batches(int64_t need) {
    for (size_t i = 0; i < batches_.size() && in_use_ + need > budget_;) {
      if (batches_[i]->n_used() == 0) {
        in_use_ -= batches_[i]->bytes;
        batches_.erase(batches_.begin() + i);
      } else {
        i++;
      }
    }
  }

  Batch* open_batch(const SparseMat<double>* A, bool fp32) {
    const int64_t bytes = batch_bytes(A, fp32);
    if (in_use_ + bytes > budget_) release_idle_batches(bytes);
    if (in_use_ + bytes > budget_) return nullptr;
    auto b = std::make_unique<Batch>();
    b->A = A;
    b->fp32 = fp32;
    b->bytes = bytes;
    b->owner.assign(slots_, {nullptr, 0});
    const Map* map = A->domain_map;
    if (fp32) {
      b->op32.A = sparse_fp32_twin(A);
      b->mb32.reset(new MinresBatch<float>(b->op32, map, slots_));
    } else {
      b->op64.A = A;
      b->mb64.reset(new MinresBatch<double>(b->op64, map, slots_));
    }
    in_use_ += bytes;
    batches_.push_back(std::move(b));
    return batches_.back().get();
  }

  // first come, first served: the queue head blocks later requests only
  // while it waits for memory, not while it waits for slots of another batch
  void admit() {
    for (auto it = queue_.begin(); it != queue_.end();) {
      Request* r = *it;
      if (r->jada) {
        if (r->bytes > budget_) {  // can never fit
          fail(r, PHIST_MEM_ALLOC_FAILED);
          it = queue_.erase(it);
          continue;
        }
        if (in_use_ + r->bytes > budget_) release_idle_batches(r->bytes);
        if (in_use_ + r->bytes > budget_) break;  // wait for memory, keep the order
        try {
          r->jd = jada_start(r->A, r->fp32, r->opts, r->Q, r->R, r->ev, r->res);
          r->state = PHIST_AMD_REQ_RUNNING;
          in_use_ += r->bytes;
          running_jd_.push_back(r);
        } catch (const Error& e) {
          fail(r, e.code < 0 ? e.code : PHIST_FUNCTIONAL_ERROR);
        }
        it = queue_.erase(it);
        continue;
      }
      Batch* b = find_batch(r->A, r->fp32);
      if (!b && batch_bytes(r->A, r->fp32) > budget_) {  // can never fit
        fail(r, PHIST_MEM_ALLOC_FAILED);
        it = queue_.erase(it);
        continue;
      }
      if (!b) b = open_batch(r->A, r->fp32);
      if (!b) break;  // no memory for a new batch: keep the order
      int slot;
      while (r->next_col < (int)r->shifts.size() && (slot = b->free_slot()) >= 0) {
        const int c = r->next_col++;
        if (b->fp32)
          b->mb32->load(slot, r->B, c, (float)r->shifts[c], r->tol, r->max_iter);
        else
          b->mb64->load(slot, r->B, c, r->shifts[c], r->tol, r->max_iter);
        b->owner[slot] = {r, c};
        r->state = PHIST_AMD_REQ_RUNNING;
      }
      if (r->next_col == (int)r->shifts.size())
        it = queue_.erase(it);
      else
        ++it;
    }
  }

  void harvest(Batch& b) {
    for (int j = 0; j < slots_; j++) {
      Request* r = b.owner[j].first;
      if (!r) continue;
      const bool active = b.fp32 ? b.mb32->active(j) : b.mb64->active(j);
      if (active) continue;
      const int c = b.owner[j].second;
      int it;
      bool conv;
      if (b.fp32) {
        b.mb32->store(j, r->X, c);
        it = b.mb32->iters(j);
        conv = b.mb32->converged(j);
        b.mb32->release(j);
      } else {
        b.mb64->store(j, r->X, c);
        it = b.mb64->iters(j);
        conv = b.mb64->converged(j);
        b.mb64->release(j);
      }
      b.owner[j] = {nullptr, 0};
      r->iters = std::max(r->iters, it);
      r->all_converged = r->all_converged && conv;
      if (--r->remaining == 0) {
        r->state = r->all_converged ? PHIST_AMD_REQ_DONE : PHIST_AMD_REQ_DONE_UNCONVERGED;
        r->t_done = now_s();
      }
    }
  }

  void fail(Request* r, int code) {
    r->state = code;
    r->t_done = now_s();
  }

  int64_t budget_, in_use_ = 0;
  int slots_;
  int64_t last_id_ = 0;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs_;
  std::deque<Request*> queue_;
  std::vector<std::unique_ptr<Batch>> batches_;
  std::vector<Request*> running_jd_;
  double t_jd_ = 0.0, t_iter_ = 0.0;  // last step: time of the eigensolver steps, of one batch iteration
};

}  // namespace
}  // namespace phx

using namespace phx;

struct phist_amd_server_s {
  Server* s;
};

extern "C" {

void phist_amd_server_create(phist_amd_server_ptr* srv, int64_t hbm_budget_bytes, int slots, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_create", [&] {
    ensure_init();
    *srv = new phist_amd_server_s{new Server(hbm_budget_bytes, slots)};
  });
}

void phist_amd_server_delete(phist_amd_server_ptr srv, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_delete", [&] {
    if (!srv) return;
    sync_stream();
    delete srv->s;
    delete srv;
  });
}

void phist_amd_server_submit_solve(phist_amd_server_ptr srv, phist_Dconst_sparseMat_ptr A, int nrhs,
                                   const double shifts[], phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, double tol,
                                   int maxIter, int fp32, int64_t* id, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_submit_solve", [&] {
    PHX_REQUIRE(srv, PHIST_INVALID_INPUT, "null server");
    *id = srv->s->submit_solve(reinterpret_cast<const SparseMat<double>*>(A), nrhs, shifts,
                               reinterpret_cast<const Mvec<double>*>(B), reinterpret_cast<Mvec<double>*>(X), tol,
                               maxIter, fp32 != 0);
  });
}

void phist_amd_server_submit_jada(phist_amd_server_ptr srv, phist_Dconst_sparseMat_ptr A, const phist_jadaOpts* opts,
                                  int fp32, phist_Dmvec_ptr Q, phist_DsdMat_ptr R, double* ev, double* resNorm,
                                  int64_t* id, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_submit_jada", [&] {
    PHX_REQUIRE(srv && opts, PHIST_INVALID_INPUT, "null argument");
    *id = srv->s->submit_jada(reinterpret_cast<const SparseMat<double>*>(A), *opts, fp32 != 0,
                              reinterpret_cast<Mvec<double>*>(Q), reinterpret_cast<SdMat<double>*>(R), ev, resNorm);
  });
}

void phist_amd_server_step(phist_amd_server_ptr srv, int* pending, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_step", [&] {
    PHX_REQUIRE(srv, PHIST_INVALID_INPUT, "null server");
    *pending = srv->s->step();
  });
}

void phist_amd_server_status(phist_amd_server_ptr srv, int64_t id, int* state, double* latency_s, int* iters,
                             int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_status", [&] {
    PHX_REQUIRE(srv, PHIST_INVALID_INPUT, "null server");
    srv->s->status(id, state, latency_s, iters);
  });
}

void phist_amd_server_cancel(phist_amd_server_ptr srv, int64_t id, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_cancel", [&] {
    PHX_REQUIRE(srv, PHIST_INVALID_INPUT, "null server");
    srv->s->cancel(id);
  });
}

void phist_amd_server_release(phist_amd_server_ptr srv, int64_t id, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_server_release", [&] {
    PHX_REQUIRE(srv, PHIST_INVALID_INPUT, "null server");
    srv->s->release(id);
  });
}

}  // extern "C"
