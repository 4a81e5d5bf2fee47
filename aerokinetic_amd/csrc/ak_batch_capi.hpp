// ak_batch_capi.hpp -- C ABI of a batch of two-level cases stepped together (included by
// ak_capi.hip after the refined ABI).
//
// A batch holds refined cases of one storage type, each with its own buffers, step count
// and history.  One batch step = one coarse step of every member, as the same kernel
// sequence as a single case, where every kernel runs for ALL members in one launch over
// their concatenated work lists (ak_batch.hpp):
//   coarse sweeps -> [coarse face flux, legacy farfield, donor moments at t + dt]
//   -> prolongation x pass of both substeps -> 2 x [prolongation y/z -> fine sweeps
//   -> [fine face flux, walls + load partials]] -> reflux -> [fine moments + restriction,
//   donor moments at t of the next step] -> loads
// (brackets: one launch; a batch run call starts with the donor moments of every member).
// The launch count per step is independent of the number of members (12), so small
// cases that are latency-bound alone fill the GPU together.  The step is captured on the
// batch stream as a HIP graph per launch variant (lead member's coarse parity x sampled)
// and replayed while the membership stays the same.

namespace {

struct Batch {
  int dev = 0, storage = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<Refined *> members;
  unsigned long long gen = 1;       // membership generation
  BatchMember *d_desc = nullptr;    // [4][cap]: coarse sweep (lead parity 0, 1), fine sweep (substep 0, 1)
  CaseDev *d_cases = nullptr;       // [cap]
  int *d_first = nullptr;           // [BP_COUNT][cap + 1] first block of each member per kernel
  uint16_t *d_bmem = nullptr;       // member of every sweep block: [nb_coarse] coarse, then [nb_fine] fine
  size_t bmem_cap = 0;
  int cap = 0, nb_coarse = 0, nb_fine = 0;
  int nblk[BP_COUNT] = {};
  size_t yz_lds = 0;                // dynamic LDS of the y/z prolongation (largest member)
  unsigned long long desc_key = 0;  // state the descriptors were built for
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  bool use_graphs = true;
  GraphCache graphs;  // slot = lead member's coarse parity * 2 + sampled
  int sweep_grid = 0;  // batched sweep: 0 one block per tile, else a persistent grid of this many blocks
};

Batch *BT(void *h) { return (Batch *)h; }

int bfail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return -1;
}

#define BCHK(call)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) return bfail("%s: %s", #call, hipGetErrorString(e_));       \
  } while (0)

// Generation of the batch's membership and of every member's kernel arguments.
unsigned long long batch_key(const Batch *b) {
  unsigned long long k = b->gen * 0x9E3779B97F4A7C15ull;
  for (const Refined *r : b->members) k = (k ^ (r->c->gen * 1315423911ull + r->f->gen)) * 0x100000001B3ull;
  return k;
}

// batched sweeps run 256-thread blocks: a member lattice's own tile when it has 256 threads
// (the descriptors are rebuilt when a member's tile changes: it bumps the lattice's
// generation, which is part of batch_key), else its default geometry
SweepGeom tile_geom(const Lattice *s) {
  return s->geom.bx * s->geom.by == 256 ? s->geom : default_geom(s->nx, s->ny, s->nz);
}

BatchMember member_of(const Lattice *s, bool odd, int first) {
  BatchMember m{};
  m.A = s->pop;
  m.solid = s->d_solid;
  m.L = s->L;
  m.G = tile_geom(s);
  m.ws = s->ws;
  m.wb = s->wb;
  m.scale = s->scale;
  m.inv_scale = s->inv_scale;
  m.diag = s->d_diag;
  m.first_block = first;
  m.odd = odd ? 1 : 0;
  return m;
}

// Descriptor tables for the current membership.  Member m's coarse parity relative to the
// lead member (members[0]) is fixed while the membership is, so two coarse tables cover
// every step; the fine parity of substep k is the same every coarse step.
int build_descriptors(Batch *b) {
  const unsigned long long key = batch_key(b);
  if (key == b->desc_key && b->d_desc) return 0;
  const int n = (int)b->members.size();
  if (n > b->cap) {
    for (void *p : {(void *)b->d_desc, (void *)b->d_cases, (void *)b->d_first})
      if (p) (void)hipFree(p);
    b->d_desc = nullptr;
    b->d_cases = nullptr;
    b->d_first = nullptr;
    b->cap = std::max(n, 2 * b->cap);
    BCHK(hipMalloc((void **)&b->d_desc, sizeof(BatchMember) * 4 * (size_t)b->cap));
    BCHK(hipMalloc((void **)&b->d_cases, sizeof(CaseDev) * (size_t)b->cap));
    BCHK(hipMalloc((void **)&b->d_first, sizeof(int) * BP_COUNT * (size_t)(b->cap + 1)));
  }
  std::vector<BatchMember> h(4 * (size_t)b->cap);
  const long long lead = b->members[0]->c->sweeps;
  for (int p = 0; p < 2; ++p) {
    int first = 0;
    for (int i = 0; i < n; ++i) {
      const Lattice *c = b->members[i]->c;
      // member sweeps when the lead member's sweep count has parity p
      const long long sw = c->sweeps + ((p - (int)(lead % 2) + 2) % 2);
      h[(size_t)p * b->cap + i] = member_of(c, sw % 2 == 0, first);
      first += tile_geom(c).nblocks;
    }
    b->nb_coarse = first;
  }
  for (int sub = 0; sub < 2; ++sub) {
    int first = 0;
    for (int i = 0; i < n; ++i) {
      const Lattice *f = b->members[i]->f;
      h[(size_t)(2 + sub) * b->cap + i] = member_of(f, (f->sweeps + sub) % 2 == 0, first);
      first += tile_geom(f).nblocks;
    }
    b->nb_fine = first;
  }
  BCHK(hipMemcpy(b->d_desc, h.data(), sizeof(BatchMember) * h.size(), hipMemcpyHostToDevice));
  // the member of every sweep block (one load per block instead of a bisection over the
  // descriptors): the coarse sweeps' blocks, then the fine sweeps'
  if (n > 65535) return bfail("%d members: at most 65535 in a batch", n);
  std::vector<uint16_t> bm((size_t)b->nb_coarse + b->nb_fine);
  for (int lev = 0; lev < 2; ++lev) {
    size_t k = lev ? (size_t)b->nb_coarse : 0;
    for (int i = 0; i < n; ++i) {
      const Lattice *l = lev ? b->members[i]->f : b->members[i]->c;
      for (int j = 0; j < tile_geom(l).nblocks; ++j) bm[k++] = (uint16_t)i;
    }
  }
  if (bm.size() > b->bmem_cap) {
    if (b->d_bmem) (void)hipFree(b->d_bmem);
    b->d_bmem = nullptr;
    b->bmem_cap = 0;
    BCHK(hipMalloc((void **)&b->d_bmem, sizeof(uint16_t) * bm.size()));
    b->bmem_cap = bm.size();
  }
  BCHK(hipMemcpy(b->d_bmem, bm.data(), sizeof(uint16_t) * bm.size(), hipMemcpyHostToDevice));
  // per-member arguments of the step kernels and their work-list tables
  std::vector<CaseDev> cd((size_t)b->cap);
  std::vector<int> first((size_t)BP_COUNT * (b->cap + 1), 0);
  for (int i = 0; i < n; ++i) {
    const Refined *r = b->members[i];
    const Lattice *c = r->c, *f = r->f;
    CaseDev &d = cd[i];
    d.Ac = c->pop; d.Af = f->pop; d.Lc = c->L; d.Lf = f->L;
    d.c_scale = c->scale; d.c_inv = c->inv_scale; d.f_scale = f->scale; d.f_inv = f->inv_scale;
    d.B = r->B; d.V = r->V; d.F = r->F; d.T = r->d_T;
    d.before = r->d_before; d.after = r->d_after; d.regc = r->d_regc; d.regf = r->d_regf; d.part = r->d_part;
    d.px = r->d_px; d.err = r->d_err;
    d.lpx = r->d_lpx; d.cols = r->d_cols;
    d.npx = r->n_lpx; d.ncols = r->n_cols;
    d.volc = r->dxc * r->dxc * r->dxc;
    d.volf = r->dxf * r->dxf * r->dxf;
    d.inv_volc = 1.0 / d.volc;
    d.loads_scale = d.volf * (r->dxf / r->dtf) / r->dtc / (0.5 * r->span);
    d.W = WallLinks{f->d_wxy, f->d_wdir, f->d_wq, f->d_welem, f->nlinks, f->nelem, f->rx_off, f->ry_off, r->dxf};
    d.wall_blocks = f->nlinks ? r->wall_blocks : 0;
    d.K = LegacyLinks{c->d_legx, c->d_legz, c->d_legdir, c->d_legkind, c->d_pdf, c->d_pdfnd, c->nleg};
    d.U = c->U;
    d.ctl = c->d_ctl;
    d.hist = c->d_hist;
    d.cpar = (int)(((c->sweeps - lead) % 2 + 2) % 2);
    d.fpar = (int)(f->sweeps % 2);
    const long long nb0 = r->B.nb[0], nb1 = r->B.nb[1], nb2 = r->B.nb[2];
    const long long items[BP_COUNT] = {
        (nb0 + 6) * (nb1 + 6) * (nb2 + 6),                 // donor moments (256)
        4LL * r->nt * c->nz * FLUX_LANES,                  // coarse face flux (256)
        (long long)c->nleg,                                // legacy farfield (256)
        (long long)r->n_lpx,                               // prolongation x work items (64)
        (long long)blocks_for(r->n_cols, PZ_COLS) * 64,    // prolongation y/z: a block per column group (64)
        4LL * (2 * r->nt) * f->nz * FLUX_LANES,            // fine face flux (256)
        (long long)d.wall_blocks * 256,                    // walls (256)
        4LL * r->nt * c->nz * 64,                          // reflux, a wave per cell (256)
        (long long)restrict_tiles(r->B) * 256};            // restriction, a block per tile (256)
    const int threads[BP_COUNT] = {256, 256, 256, 64, 64, 256, 256, 256, 256};
    for (int k = 0; k < BP_COUNT; ++k)
      first[(size_t)k * (b->cap + 1) + i + 1] = first[(size_t)k * (b->cap + 1) + i] + (int)blocks_for(items[k], threads[k]);
  }
  for (int k = 0; k < BP_COUNT; ++k) b->nblk[k] = first[(size_t)k * (b->cap + 1) + n];
  b->yz_lds = 0;
  for (const Refined *r : b->members) b->yz_lds = std::max(b->yz_lds, prolong_yz_lds(r->B, r->c->storage == 0 ? sizeof(double) : sizeof(float)));
  BCHK(hipMemcpy(b->d_cases, cd.data(), sizeof(CaseDev) * cd.size(), hipMemcpyHostToDevice));
  BCHK(hipMemcpy(b->d_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice));
  b->desc_key = key;
  return 0;
}

template <int S>
void batched_sweep(Batch *b, const BatchMember *d, const uint16_t *bmem, int nblocks) {
  const int grid = b->sweep_grid > 0 ? std::min(nblocks, b->sweep_grid) : nblocks;
  hipLaunchKernelGGL((k_sweep_batch<S, false>), dim3(grid), dim3(256), 0, b->stream, d, bmem, nblocks);
}

// Donor moments of every member's current state at the start of a batch run (each step's
// restriction launch then leaves those of the next step).
template <int S>
void batch_start(Batch *b) {
  const int nm = (int)b->members.size(), lead = (int)(b->members[0]->c->sweeps % 2);
  if (b->nblk[BP_DONOR] > 0)
    hipLaunchKernelGGL((kb_donor<S>), dim3(b->nblk[BP_DONOR]), dim3(256), 0, b->stream, b->d_cases,
                       b->d_first + (size_t)BP_DONOR * (b->cap + 1), nm, lead, 0, 0);
}

// One coarse step of every member: the single-case kernel sequence, each kernel one launch
// over all members (host counters advance as in refined_coarse_step).
template <int S>
void batch_step(Batch *b, bool sample) {
  const int nm = (int)b->members.size(), lead = (int)(b->members[0]->c->sweeps % 2), nc = b->cap;
  hipStream_t st = b->stream;
  const CaseDev *C = b->d_cases;
  auto F = [b](int k) { return b->d_first + (size_t)k * (b->cap + 1); };
  auto go = [&](int k) { return b->nblk[k] > 0; };
  batched_sweep<S>(b, b->d_desc + (size_t)lead * nc, b->d_bmem, b->nb_coarse);
  for (Refined *r : b->members) r->c->sweeps++;
  // face flux, farfield and donor moments of the new coarse state: one launch (disjoint slots)
  {
    const int nfl = b->nblk[BP_FLUXC], nlg = b->nblk[BP_LEGACY], ndn = b->nblk[BP_DONOR];
    if (nfl + nlg + ndn > 0)
      hipLaunchKernelGGL((kb_after_coarse<S>), dim3(nfl + nlg + ndn), dim3(256), 0, st, C, F(BP_FLUXC), F(BP_LEGACY),
                         F(BP_DONOR), nm, lead, nfl, nlg);
  }
  // x pass of the prolongation for both substeps
  if (go(BP_PX)) hipLaunchKernelGGL((kb_prolong_x<S>), dim3(b->nblk[BP_PX]), dim3(64), 0, st, C, F(BP_PX), nm);
  for (int sub = 0; sub < 2; ++sub) {
    if (go(BP_PYZ)) hipLaunchKernelGGL((kb_prolong_yz<S>), dim3(b->nblk[BP_PYZ]), dim3(64), b->yz_lds, st, C, F(BP_PYZ), nm, sub);
    batched_sweep<S>(b, b->d_desc + (size_t)(2 + sub) * nc, b->d_bmem + b->nb_coarse, b->nb_fine);
    for (Refined *r : b->members) r->f->sweeps++;
    // fine face flux and walls + loads: one launch
    if (b->nblk[BP_FLUXF] + b->nblk[BP_WALLS] > 0)
      hipLaunchKernelGGL((kb_after_fine<S>), dim3(b->nblk[BP_FLUXF] + b->nblk[BP_WALLS]), dim3(256), 0, st, C,
                         F(BP_FLUXF), F(BP_WALLS), nm, sub + 1, sub, b->nblk[BP_FLUXF]);
  }
  if (go(BP_REFLUX)) hipLaunchKernelGGL((kb_reflux<S>), dim3(b->nblk[BP_REFLUX]), dim3(256), 0, st, C, F(BP_REFLUX), nm, lead, 1);
  // restriction + the next step's donor moments: one launch
  if (b->nblk[BP_RESTRICT] + b->nblk[BP_DONOR] > 0)
    hipLaunchKernelGGL((kb_restrict<S>), dim3(b->nblk[BP_RESTRICT] + b->nblk[BP_DONOR]), dim3(256), 0, st, C,
                       F(BP_RESTRICT), F(BP_DONOR), nm, lead, b->nblk[BP_RESTRICT]);
  hipLaunchKernelGGL(kb_loads, dim3(nm), dim3(64), 0, st, C, sample ? 1 : 0);
}

// Order the batch stream after all work queued on the members' streams (or the reverse).
int order_after_members(Batch *b) {
  for (size_t i = 0; i < b->members.size(); ++i) {
    BCHK(hipEventRecord(b->join[i], b->members[i]->c->stream));
    BCHK(hipStreamWaitEvent(b->stream, b->join[i], 0));
  }
  return 0;
}
int order_members_after(Batch *b) {
  BCHK(hipEventRecord(b->fork, b->stream));
  for (Refined *r : b->members) BCHK(hipStreamWaitEvent(r->c->stream, b->fork, 0));
  return 0;
}

}  // namespace

extern "C" {

int ak_batch_create(void **handle, int device, int storage, void *stream) {
  if (!handle) return bfail("null handle pointer");
  *handle = nullptr;
  if (storage < 0 || storage > 2) return bfail("unknown storage precision %d", storage);
  if (hipSetDevice(device) != hipSuccess) return bfail("hipSetDevice(%d) failed", device);
  Batch *b = new Batch();
  b->dev = device;
  b->storage = storage;
  b->stream = (hipStream_t)stream;
  if (!b->stream) {
    // non-blocking: cases being set up on the null stream while a chunk runs do not wait for it
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
      delete b;
      return bfail("hipStreamCreate failed");
    }
    b->own_stream = true;
  }
  if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return bfail("hipEventCreate failed");
  }
  *handle = b;
  return 0;
}

int ak_batch_destroy(void *handle) {
  Batch *b = BT(handle);
  if (!b) return 0;
  (void)hipSetDevice(b->dev);
  (void)hipStreamSynchronize(b->stream);
  b->graphs.clear();
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  (void)hipEventDestroy(b->fork);
  for (void *p : {(void *)b->d_desc, (void *)b->d_cases, (void *)b->d_first, (void *)b->d_bmem})
    if (p) (void)hipFree(p);
  if (b->own_stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int ak_batch_add(void *handle, void *refined) {
  Batch *b = BT(handle);
  Refined *r = R(refined);
  if (!b || !r) return bfail("null batch or refined handle");
  if (r->c->storage != b->storage) return bfail("member storage %d differs from the batch's %d", r->c->storage, b->storage);
  if (r->c->dev != b->dev) return bfail("member on device %d, batch on device %d", r->c->dev, b->dev);
  if (check_ready(r->c) || check_ready(r->f)) return bfail("member populations not bound");
  for (Refined *m : b->members)
    if (m == r) return bfail("case already in the batch");
  if (r->f->sweeps % 2 != 1) return bfail("member not at a coarse-step boundary");
  hipEvent_t e = nullptr;
  BCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  BCHK(hipStreamSynchronize(b->stream));
  b->graphs.clear();
  b->join.push_back(e);
  b->members.push_back(r);
  b->gen++;
  return (int)b->members.size() - 1;
}

int ak_batch_remove(void *handle, void *refined) {
  Batch *b = BT(handle);
  if (!b) return bfail("null batch handle");
  auto it = std::find(b->members.begin(), b->members.end(), R(refined));
  if (it == b->members.end()) return bfail("case not in the batch");
  // the batch's work is finished and its graphs are gone before the member can be used
  // alone again or destroyed
  BCHK(hipStreamSynchronize(b->stream));
  b->graphs.clear();
  if (order_members_after(b)) return -1;
  b->members.erase(it);
  (void)hipEventDestroy(b->join.back());
  b->join.pop_back();
  b->gen++;
  return 0;
}

int ak_batch_size(void *handle) { return handle ? (int)BT(handle)->members.size() : 0; }

int ak_batch_set_graph_mode(void *handle, int enable) {
  Batch *b = BT(handle);
  if (!b) return bfail("null batch handle");
  b->use_graphs = enable != 0;
  return 0;
}

int ak_batch_set_sweep_grid(void *handle, int blocks) {
  Batch *b = BT(handle);
  if (!b) return bfail("null batch handle");
  if (blocks < 0) return bfail("sweep grid of %d blocks", blocks);
  b->sweep_grid = blocks == 0 ? 0 : std::max(8, blocks / 8 * 8);
  b->gen++;  // the captured graphs hold the old launch shape
  return 0;
}

// nsteps coarse steps of every member, enqueued on the batch stream (returns before they
// run).  sample_every is 1 (loads every step, to each member's history) or 0 (none).  Step
// numbers continue each member's own count.  Synchronise the batch (ak_batch_synchronize)
// before reading a member's history or state.
int ak_batch_run(void *handle, int nsteps, int sample_every) {
  Batch *b = BT(handle);
  if (!b) return bfail("null batch handle");
  if (sample_every != 0 && sample_every != 1) return bfail("batch runs sample every step (1) or never (0)");
  if (b->members.empty() || nsteps <= 0) return 0;
  BCHK(hipSetDevice(b->dev));
  if (build_descriptors(b)) return -1;
  if (order_after_members(b)) return -1;
  const bool sample = sample_every == 1;
  for (Refined *r : b->members) {
    set_ctl_on(r->c, b->stream, r->steps + 1, r->c->hist_rows);
    r->budget_complete = false;  // batched steps keep no conservation budget
  }
  switch (b->storage) {
    case 0: batch_start<0>(b); break;
    case 1: batch_start<1>(b); break;
    default: batch_start<2>(b); break;
  }
  for (int k = 0; k < nsteps; ++k) {
    if (sample) {
      bool full = false;
      for (Refined *r : b->members) full |= r->c->hist_rows >= r->c->hist_cap;
      if (full) {
        BCHK(hipStreamSynchronize(b->stream));
        for (Refined *r : b->members) {
          if (r->c->hist_rows < r->c->hist_cap) continue;
          if (flush_history(r->c)) return bfail("history flush: %s", r->c->err.c_str());
          set_ctl_on(r->c, b->stream, r->steps + 1, 0);
        }
      }
    }
    auto launch = [&] {
      switch (b->storage) {
        case 0: batch_step<0>(b, sample); break;
        case 1: batch_step<1>(b, sample); break;
        default: batch_step<2>(b, sample); break;
      }
    };
    auto advance = [&] {
      for (Refined *r : b->members) {
        r->c->sweeps += 1;
        r->f->sweeps += 2;
      }
    };
    if (b->use_graphs) {
      const int slot = (int)(b->members[0]->c->sweeps % 2) * 2 + (sample ? 1 : 0);
      if (graph_step(b->members[0]->c, b->stream, b->graphs, slot, batch_key(b), launch, advance))
        return bfail("batch graph: %s", b->members[0]->c->err.c_str());
    } else {
      launch();
    }
    for (Refined *r : b->members) {
      if (sample) r->c->hist_rows++;
      r->steps++;
    }
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bfail("batch run: %s", hipGetErrorString(e));
  return 0;
}

int ak_batch_synchronize(void *handle) {
  Batch *b = BT(handle);
  if (!b) return bfail("null batch handle");
  BCHK(hipSetDevice(b->dev));
  BCHK(hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"
