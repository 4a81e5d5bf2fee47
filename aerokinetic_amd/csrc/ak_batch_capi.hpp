// ak_batch_capi.hpp -- C ABI of a batch of two-level cases stepped together (included by
// ak_capi.hip after the refined ABI).
//
// A batch holds refined cases of one storage type, each with its own buffers, step count,
// history and stream.  One batch step = one coarse step of every member:
//   members' donor moments -> ONE batched coarse sweep over all members' tiles ->
//   members' coarse flux / farfield / donor moments / prolongation -> ONE batched fine
//   sweep -> members' fine flux / walls / prolongation -> ONE batched fine sweep ->
//   members' fine flux / walls / reflux / restriction / loads.
// The per-member phases are forked onto the members' own streams and joined before each
// batched sweep, so small transfer kernels of different cases overlap on the device.  The
// whole batch step is captured once per launch variant as a HIP graph and replayed while
// the membership (and every member's kernel arguments) stays the same.

namespace {

struct Batch {
  int dev = 0, storage = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<Refined *> members;
  unsigned long long gen = 1;       // membership generation
  BatchMember *d_desc = nullptr;    // [4][cap]: coarse sweep (lead parity 0, 1), fine sweep (substep 0, 1)
  int cap = 0, nb_coarse = 0, nb_fine = 0;
  unsigned long long desc_key = 0;  // state the descriptors were built for
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  bool use_graphs = true;
  GraphCache graphs;  // slot = lead member's coarse parity * 2 + sampled
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

SweepGeom tile_geom(const Lattice *s) {
  SweepGeom G{};
  G.bx = 64;
  G.by = 4;
  G.gx = (s->nx + 63) / 64;
  G.gy = (s->ny + 3) / 4;
  G.gz = s->nz;
  G.nblocks = G.gx * G.gy * G.gz;
  return G;
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
    if (b->d_desc) (void)hipFree(b->d_desc);
    b->d_desc = nullptr;
    b->cap = std::max(n, 2 * b->cap);
    BCHK(hipMalloc((void **)&b->d_desc, sizeof(BatchMember) * 4 * (size_t)b->cap));
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
  b->desc_key = key;
  return 0;
}

template <int S>
void batched_sweep(Batch *b, const BatchMember *d, int nblocks) {
  hipLaunchKernelGGL((k_sweep_batch<S, false>), dim3(nblocks), dim3(256), 0, b->stream, d, (int)b->members.size(),
                     nblocks);
}

// Run `phase(member)` for every member on the member's own stream, forked from and joined
// back into the batch stream.
template <class F>
void fork_join(Batch *b, F &&phase) {
  (void)hipEventRecord(b->fork, b->stream);
  for (size_t i = 0; i < b->members.size(); ++i) {
    Refined *r = b->members[i];
    (void)hipStreamWaitEvent(r->c->stream, b->fork, 0);
    phase(r);
    (void)hipEventRecord(b->join[i], r->c->stream);
  }
  for (size_t i = 0; i < b->members.size(); ++i) (void)hipStreamWaitEvent(b->stream, b->join[i], 0);
}

// One coarse step of every member (host counters advance as in refined_coarse_step).
template <int S>
void batch_step(Batch *b, bool sample) {
  const int lead_parity = (int)(b->members[0]->c->sweeps % 2);
  const int nc = b->cap;
  fork_join(b, [](Refined *r) { phase_before_coarse<S>(r); });
  batched_sweep<S>(b, b->d_desc + (size_t)lead_parity * nc, b->nb_coarse);
  for (Refined *r : b->members) r->c->sweeps++;
  fork_join(b, [](Refined *r) {
    phase_after_coarse<S>(r);
    phase_before_fine<S>(r, 0);
  });
  batched_sweep<S>(b, b->d_desc + (size_t)2 * nc, b->nb_fine);
  for (Refined *r : b->members) r->f->sweeps++;
  fork_join(b, [](Refined *r) {
    phase_after_fine<S>(r, 0);
    phase_before_fine<S>(r, 1);
  });
  batched_sweep<S>(b, b->d_desc + (size_t)3 * nc, b->nb_fine);
  for (Refined *r : b->members) r->f->sweeps++;
  fork_join(b, [sample](Refined *r) {
    phase_after_fine<S>(r, 1);
    phase_end<S>(r, sample);
  });
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
    if (hipStreamCreate(&b->stream) != hipSuccess) {
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
  if (b->d_desc) (void)hipFree(b->d_desc);
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
  if (r->c->stream == b->stream) return bfail("a member must have a stream of its own");
  for (Refined *m : b->members) {
    if (m == r) return bfail("case already in the batch");
    if (m->c->stream == r->c->stream) return bfail("two members share a stream");
  }
  if (r->f->sweeps % 2 != 1) return bfail("member not at a coarse-step boundary");
  hipEvent_t e = nullptr;
  BCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
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
  // the member's stream must see the batch's work before it is used alone again
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

// nsteps coarse steps of every member.  sample_every is 1 (loads every step, to each
// member's history) or 0 (none).  Step numbers continue each member's own count.
int ak_batch_run(void *handle, int nsteps, int sample_every) {
  Batch *b = BT(handle);
  if (!b) return bfail("null batch handle");
  if (sample_every != 0 && sample_every != 1) return bfail("batch runs sample every step (1) or never (0)");
  if (b->members.empty() || nsteps <= 0) return 0;
  BCHK(hipSetDevice(b->dev));
  if (build_descriptors(b)) return -1;
  if (order_after_members(b)) return -1;
  const bool sample = sample_every == 1;
  for (Refined *r : b->members) set_ctl_on(r->c, b->stream, r->steps + 1, r->c->hist_rows);
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
  if (order_members_after(b)) return -1;
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
