// eph_checkpoint.cpp -- full-precision binary checkpoints of the fix's state
// and of the harness system around it (eph_amd.h, eph_md.h).
//
// The reference keeps only the T_e grid across runs, as text with six
// significant digits (EPH_FDM::save_state, eph_fdm.h:226-265), so a resumed
// run is not the run that went on.  These files hold the binary doubles, so
// that loading one and going on gives the bits of the run that was saved.
//
// Layout (native endianness):
//   char magic[8] "EPHCKPT1"; int32 kind (1 fix, 2 harness and fix),
//   colour, atomic, model, flags, 0;
//   int64 step; float64 e0, e1 (E_local and Ee, or Ee and T_e for eph/atomic);
//   int32 nx, ny, nz, 0; float64 T[nx ny nz]                 (no grid: 0 0 0)
//   int32 n, ncols; int32 tag[n]; float64 state[n][ncols]    (per local atom:
//     ncols 6 = the two filter states of eph/coloured/exp*, 1 = E_i of
//     eph/atomic, 0 otherwise)
//   kind 2 then: int32 nlocal, nghost, ghost_cap, cap_f, cap_l, stride,
//   fstride, 0; float64x4 pos[nall_cap], vel[nall_cap]; int32 type, mask,
//   tag [nall_cap]; float64x4 shift[ghost_cap]; int32 owner[ghost_cap];
//   float64x4 xref[nlocal]; float64 f[3 fstride]; int32 nbr[cap_f stride],
//   nnbr[stride], lj[cap_l stride], ljn[stride]  (nall_cap = nlocal + ghost_cap)
//   and, when the last int32 of that header is 1, the density sweep's
//   candidate lists: int64 sizes n_cand, n_ncand, n_cref; int32 state[8],
//   ncand[n_ncand], cand[n_cand]; float64x4 cref[n_cref].  They are what the
//   lists were when the checkpoint was taken, so a restored run takes the
//   same kernel paths (candidate walk or full list) as the run that went on;
//   without them the first density sweep after a load takes the full list.
#include <cstdio>
#include <cstring>
#include <memory>
#include <unordered_map>
#include <vector>

#include "eph_harness.hpp"

extern "C" {
#include "eph_md.h"
}

using namespace eph;

namespace {

constexpr char kMagic[8] = {'E', 'P', 'H', 'C', 'K', 'P', 'T', '1'};

// a checkpoint file, or the harness's in-memory checkpoint (mem)
struct File {
  std::FILE* f = nullptr;
  std::vector<char>* mem = nullptr;
  size_t at = 0;
  bool ok = true;
  ~File() {
    if (f) std::fclose(f);
  }
  template <typename T>
  void put(const T* p, size_t n) {
    if (!ok || !n) return;
    if (mem) {
      const char* b = reinterpret_cast<const char*>(p);
      mem->insert(mem->end(), b, b + n * sizeof(T));
    } else if (std::fwrite(p, sizeof(T), n, f) != n) {
      ok = false;
    }
  }
  template <typename T>
  void get(T* p, size_t n) {
    if (!ok || !n) return;
    if (mem) {
      if (at + n * sizeof(T) > mem->size()) { ok = false; return; }
      std::memcpy(p, mem->data() + at, n * sizeof(T));
      at += n * sizeof(T);
    } else if (std::fread(p, sizeof(T), n, f) != n) {
      ok = false;
    }
  }
  long tell() const { return mem ? static_cast<long>(at) : std::ftell(f); }
  // bytes after the current position (-1 if unknown): sizes read from a file
  // are checked against it before anything is allocated for them
  long remaining() {
    if (mem) return static_cast<long>(mem->size() - at);
    const long here = std::ftell(f);
    if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long end = std::ftell(f);
    if (std::fseek(f, here, SEEK_SET) != 0 || end < here) return -1;
    return end - here;
  }
  bool seek(long off, int whence) {
    if (!mem) return std::fseek(f, off, whence) == 0;
    const long base = whence == SEEK_SET ? 0 : static_cast<long>(at);
    if (base + off < 0 || static_cast<size_t>(base + off) > mem->size()) return false;
    at = static_cast<size_t>(base + off);
    return true;
  }
};

// device -> file through a host buffer (the context's stream)
template <typename T>
int put_dev(eph_context* c, File& fl, const T* src, size_t n) {
  std::vector<T> h(n);
  if (n) {
    EPH_HIP(c, hipMemcpyAsync(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    EPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  fl.put(h.data(), n);
  return EPH_OK;
}

template <typename T>
int get_dev(eph_context* c, File& fl, T* dst, size_t n) {
  std::vector<T> h(n);
  fl.get(h.data(), n);
  if (!fl.ok) return c->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
  if (n) {
    EPH_HIP(c, hipMemcpyAsync(dst, h.data(), n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    EPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  return EPH_OK;
}

int state_cols(const eph_context* c) { return c->atomic() ? 1 : c->colour >= 2 ? 6 : 0; }

int save_fix(eph_context* c, File& fl, int kind) {
  const int hdr[6] = {kind, c->colour, c->atomic() ? 1 : 0, c->model, c->flags, 0};
  fl.put(kMagic, 8);
  fl.put(hdr, 6);
  const long long step = c->step;
  fl.put(&step, 1);
  double e[2] = {0.0, 0.0};
  if (c->atomic()) {
    e[0] = c->at->Ee;
    e[1] = c->at->Te;
  } else {
    EPH_HIP(c, hipMemcpyAsync(e, c->scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  fl.put(e, 2);
  int dims[4] = {0, 0, 0, 0};
  if (!c->atomic()) {
    dims[0] = (int)c->grid->h.nx; dims[1] = (int)c->grid->h.ny; dims[2] = (int)c->grid->h.nz;
  }
  fl.put(dims, 4);
  if (!c->atomic()) {
    if (int rc = put_dev(c, fl, c->grid->T.p, c->grid->h.ntotal())) return rc;
  }
  const int n = c->have_atoms ? c->nlocal : 0, nc = state_cols(c);
  const int sz[2] = {n, nc};
  fl.put(sz, 2);
  if (int rc = put_dev(c, fl, c->tag.p, n)) return rc;
  std::vector<double> st(static_cast<size_t>(n) * nc), col(3 * static_cast<size_t>(n));
  if (n && nc == 6) {
    for (int which = 0; which < 2; ++which) {
      if (int rc = eph_get_filter_state(c, which, col.data())) return rc;
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) st[6 * (size_t)i + 3 * which + d] = col[3 * (size_t)i + d];
    }
  } else if (n && nc == 1) {
    if (int rc = eph_get_atomic_energy(c, st.data())) return rc;
  }
  fl.put(st.data(), st.size());
  return fl.ok ? EPH_OK : c->fail(EPH_ERR_FILE, "cannot write the checkpoint file");
}

// the header and the fix's state; the per-atom rows go to the context's
// locals by tag (every local must be found).  Everything is read and checked
// before the fix is changed, so a refused file leaves the fix as it was.
int load_fix(eph_context* c, File& fl, int want_kind) {
  char magic[8];
  int hdr[6];
  fl.get(magic, 8);
  fl.get(hdr, 6);
  if (!fl.ok || std::memcmp(magic, kMagic, 8) != 0) return c->fail(EPH_ERR_FORMAT, "not an eph checkpoint file");
  if (hdr[0] != want_kind)
    return c->fail(EPH_ERR_FORMAT, want_kind == 1 ? "a harness checkpoint: load it with eph_md_checkpoint_load"
                                                  : "a fix checkpoint holds no harness system");
  if (hdr[1] != c->colour || hdr[2] != (c->atomic() ? 1 : 0) || hdr[3] != c->model || hdr[4] != c->flags)
    return c->fail(EPH_ERR_FORMAT, "checkpoint of another fix style, model or flags");
  long long step = 0;
  double e[2];
  int dims[4];
  fl.get(&step, 1);
  fl.get(e, 2);
  fl.get(dims, 4);
  if (!fl.ok) return c->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
  std::vector<double> T;
  if (c->atomic()) {
    if (dims[0] || dims[1] || dims[2]) return c->fail(EPH_ERR_FORMAT, "checkpoint grid does not match");
  } else {
    if (dims[0] != (int)c->grid->h.nx || dims[1] != (int)c->grid->h.ny || dims[2] != (int)c->grid->h.nz)
      return c->fail(EPH_ERR_FORMAT, "checkpoint grid does not match the fix's T_e grid");
    T.resize(c->grid->h.ntotal());
    fl.get(T.data(), T.size());
  }
  int sz[2] = {-1, -1};
  fl.get(sz, 2);
  if (!fl.ok) return c->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
  if (sz[0] < 0 || sz[1] != state_cols(c)) return c->fail(EPH_ERR_FORMAT, "checkpoint per-atom block does not match");
  const long left = fl.remaining();
  if (left < 0 || static_cast<double>(sz[0]) * (sizeof(int) + sizeof(double) * sz[1]) > static_cast<double>(left))
    return c->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
  std::vector<int> tags(sz[0]);
  std::vector<double> st(static_cast<size_t>(sz[0]) * sz[1]);
  fl.get(tags.data(), tags.size());
  fl.get(st.data(), st.size());
  if (!fl.ok) return c->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
  // the per-atom state in the context's local order
  const int nc = sz[1];
  const int n = (nc > 0 && c->have_atoms) ? c->nlocal : 0;
  if (nc > 0 && !c->have_atoms) return c->fail(EPH_ERR_STATE, "set the atoms before loading per-atom state");
  std::vector<double> col(nc == 6 ? 6 * static_cast<size_t>(n) : 0), E(nc == 1 ? n : 0);
  if (n) {
    std::vector<int> mine(n);
    EPH_HIP(c, hipMemcpyAsync(mine.data(), c->tag.p, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EPH_HIP(c, hipStreamSynchronize(c->stream));
    std::unordered_map<int, int> row;
    row.reserve(tags.size() * 2);
    for (size_t r = 0; r < tags.size(); ++r) row[tags[r]] = (int)r;
    for (int i = 0; i < n; ++i) {
      const auto it = row.find(mine[i]);
      if (it == row.end()) return c->fail(EPH_ERR_FORMAT, "an atom's tag is not in the checkpoint");
      const double* s = &st[(size_t)it->second * nc];
      if (nc == 1) {
        E[i] = s[0];
      } else {
        for (int which = 0; which < 2; ++which)
          for (int d = 0; d < 3; ++d) col[3 * ((size_t)which * n + i) + d] = s[3 * which + d];
      }
    }
  }
  // apply
  if (!c->atomic()) {
    EPH_HIP(c, hipMemcpyAsync(c->grid->T.p, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EPH_HIP(c, hipMemcpyAsync(c->scal.p, e, 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EPH_HIP(c, hipStreamSynchronize(c->stream));
  }
  if (nc == 6 && n) {
    for (int which = 0; which < 2; ++which)
      if (int rc = eph_set_filter_state(c, which, col.data() + 3 * (size_t)which * n)) return rc;
  } else if (nc == 1 && n) {
    if (int rc = eph_set_atomic_energy(c, E.data())) return rc;
  }
  if (c->atomic()) {
    c->at->Ee = e[0];
    c->at->Te = e[1];
  }
  c->step = step;
  return EPH_OK;
}

int open_file(File& fl, const char* path, const char* mode) {
  fl.f = std::fopen(path, mode);
  return fl.f ? EPH_OK : EPH_ERR_FILE;
}

}  // namespace

extern "C" {

int eph_checkpoint_save(eph_context* ctx, const char* path) {
  if (!ctx || !path) return EPH_ERR_ARGS;
  File fl;
  if (open_file(fl, path, "wb")) return ctx->fail(EPH_ERR_FILE, std::string("cannot write ") + path);
  return save_fix(ctx, fl, 1);
}

int eph_checkpoint_load(eph_context* ctx, const char* path) {
  if (!ctx || !path) return EPH_ERR_ARGS;
  File fl;
  if (open_file(fl, path, "rb")) return ctx->fail(EPH_ERR_FILE, std::string("cannot read ") + path);
  return load_fix(ctx, fl, 1);
}

int eph_md_checkpoint_save(eph_md* md, eph_context* fix, const char* path) {
  if (!md || !fix) return EPH_ERR_ARGS;
  if (fix->atomic()) return fix->fail(EPH_ERR_UNSUPPORTED, "harness checkpoints hold the eph styles with a T_e grid");
  if (!fix->have_atoms || !fix->have_list || md->ghost_cap <= 0)
    return fix->fail(EPH_ERR_STATE, "attach the fix to the harness first");
  File fl;
  if (path) {
    if (open_file(fl, path, "wb")) return fix->fail(EPH_ERR_FILE, std::string("cannot write ") + path);
  } else {
    md->ckpt.clear();
    fl.mem = &md->ckpt;
  }
  if (int rc = save_fix(fix, fl, 2)) return rc;
  const int st = fix->stride;
  const bool cands = fix->cstate.p && fix->cand.p && fix->ncand.p && fix->cref.p;
  const int hb[8] = {md->nlocal, fix->nghost, md->ghost_cap, md->cap_f, md->cap_l, st, fix->fstride,
                     cands ? 1 : 0};
  fl.put(hb, 8);
  const size_t na = (size_t)md->nlocal + md->ghost_cap, gc = md->ghost_cap;
  int rc;
  if ((rc = put_dev(fix, fl, fix->pos.p, na)) || (rc = put_dev(fix, fl, fix->vel.p, na)) ||
      (rc = put_dev(fix, fl, fix->type.p, na)) || (rc = put_dev(fix, fl, fix->mask.p, na)) ||
      (rc = put_dev(fix, fl, fix->tag.p, na)) || (rc = put_dev(fix, fl, fix->ghost_shift.p, gc)) ||
      (rc = put_dev(fix, fl, fix->ghost_owner.p, gc)) || (rc = put_dev(fix, fl, md->xref.p, md->nlocal)) ||
      (rc = put_dev(fix, fl, fix->f.p, 3 * (size_t)fix->fstride)) ||
      (rc = put_dev(fix, fl, fix->nbr.p, (size_t)md->cap_f * st)) ||
      (rc = put_dev(fix, fl, fix->nnbr.p, st)) ||
      (rc = put_dev(fix, fl, md->lj_nbr.p, (size_t)md->cap_l * st)) ||
      (rc = put_dev(fix, fl, md->lj_nn.p, st)))
    return rc;
  if (cands) {
    const long long sz[3] = {(long long)fix->cand.n, (long long)fix->ncand.n, (long long)fix->cref.n};
    fl.put(sz, 3);
    if ((rc = put_dev(fix, fl, fix->cstate.p, 8)) || (rc = put_dev(fix, fl, fix->ncand.p, fix->ncand.n)) ||
        (rc = put_dev(fix, fl, fix->cand.p, fix->cand.n)) || (rc = put_dev(fix, fl, fix->cref.p, fix->cref.n)))
      return rc;
  }
  return fl.ok ? EPH_OK : fix->fail(EPH_ERR_FILE, "cannot write the checkpoint file");
}

int eph_md_checkpoint_load(eph_md* md, eph_context* fix, const char* path) {
  if (!md || !fix) return EPH_ERR_ARGS;
  if (fix->atomic()) return fix->fail(EPH_ERR_UNSUPPORTED, "harness checkpoints hold the eph styles with a T_e grid");
  if (!fix->have_atoms || !fix->have_list || md->ghost_cap <= 0)
    return fix->fail(EPH_ERR_STATE, "attach the fix to the harness first");
  File fl;
  if (path) {
    if (open_file(fl, path, "rb")) return fix->fail(EPH_ERR_FILE, std::string("cannot read ") + path);
  } else {
    if (md->ckpt.empty()) return fix->fail(EPH_ERR_STATE, "no checkpoint kept in memory");
    fl.mem = &md->ckpt;
  }
  // the fix's block comes first in the file; its per-atom rows are matched
  // to the atoms restored below, so they are read again after those
  const long pos0 = fl.tell();
  {
    // skip to the harness block: read the fix block without applying it
    char magic[8];
    int hdr[6], dims[4], sz[2];
    long long step;
    double e[2];
    fl.get(magic, 8); fl.get(hdr, 6); fl.get(&step, 1); fl.get(e, 2); fl.get(dims, 4);
    if (!fl.ok || std::memcmp(magic, kMagic, 8) != 0) return fix->fail(EPH_ERR_FORMAT, "not an eph checkpoint file");
    if (hdr[0] != 2) return fix->fail(EPH_ERR_FORMAT, "a fix checkpoint holds no harness system");
    const long G = (long)dims[0] * dims[1] * dims[2];
    if (G < 0 || !fl.seek(G * (long)sizeof(double), SEEK_CUR)) return fix->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
    fl.get(sz, 2);
    if (!fl.ok || sz[0] < 0 || sz[1] < 0) return fix->fail(EPH_ERR_FORMAT, "checkpoint per-atom block does not match");
    const long skip = (long)sz[0] * (long)sizeof(int) + (long)sz[0] * sz[1] * (long)sizeof(double);
    if (!fl.seek(skip, SEEK_CUR)) return fix->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
  }
  int hb[8];
  fl.get(hb, 8);
  const int st = fix->stride;
  if (!fl.ok || hb[0] != md->nlocal || hb[2] != md->ghost_cap || hb[3] != md->cap_f || hb[4] != md->cap_l ||
      hb[5] != st || hb[6] != fix->fstride || hb[1] < 0 || hb[1] > md->ghost_cap)
    return fix->fail(EPH_ERR_FORMAT, "checkpoint of another harness system (atoms or capacities differ)");
  const size_t na = (size_t)md->nlocal + md->ghost_cap, gc = md->ghost_cap;
  int rc;
  if ((rc = get_dev(fix, fl, fix->pos.p, na)) || (rc = get_dev(fix, fl, fix->vel.p, na)) ||
      (rc = get_dev(fix, fl, fix->type.p, na)) || (rc = get_dev(fix, fl, fix->mask.p, na)) ||
      (rc = get_dev(fix, fl, fix->tag.p, na)) || (rc = get_dev(fix, fl, fix->ghost_shift.p, gc)) ||
      (rc = get_dev(fix, fl, fix->ghost_owner.p, gc)) || (rc = get_dev(fix, fl, md->xref.p, md->nlocal)) ||
      (rc = get_dev(fix, fl, fix->f.p, 3 * (size_t)fix->fstride)) ||
      (rc = get_dev(fix, fl, fix->nbr.p, (size_t)md->cap_f * st)) ||
      (rc = get_dev(fix, fl, fix->nnbr.p, st)) ||
      (rc = get_dev(fix, fl, md->lj_nbr.p, (size_t)md->cap_l * st)) ||
      (rc = get_dev(fix, fl, md->lj_nn.p, st)))
    return rc;
  bool cands = false;
  if (hb[7] == 1) {
    long long sz[3];
    fl.get(sz, 3);
    if (!fl.ok || sz[0] < 0 || sz[1] < 0 || sz[2] < 0) return fix->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
    cands = fix->cstate.p && fix->cand.p && (size_t)sz[0] == fix->cand.n && (size_t)sz[1] == fix->ncand.n &&
            (size_t)sz[2] == fix->cref.n;
    if (cands) {
      if ((rc = get_dev(fix, fl, fix->cstate.p, 8)) || (rc = get_dev(fix, fl, fix->ncand.p, fix->ncand.n)) ||
          (rc = get_dev(fix, fl, fix->cand.p, fix->cand.n)) || (rc = get_dev(fix, fl, fix->cref.p, fix->cref.n)))
        return rc;
    } else if (!fl.seek((long)(8 + sz[0] + sz[1]) * (long)sizeof(int) + (long)sz[2] * (long)sizeof(double4),
                        SEEK_CUR)) {
      return fix->fail(EPH_ERR_FORMAT, "checkpoint file ends early");
    }
  } else if (hb[7] != 0) {
    return fix->fail(EPH_ERR_FORMAT, "checkpoint of another harness system (atoms or capacities differ)");
  }
  fix->nghost = hb[1];
  md->nghost = hb[1];
  EPH_HIP(fix, launch_fill_int(md->flags.p, 8, 0, fix->stream));
  if (!cands) fix->drop_candidates();  // other capacities: retaken by the next density sweep
  fix->pv_fresh = false;
  // now the fix's block, with the restored atoms' tags
  if (!fl.seek(pos0, SEEK_SET)) return fix->fail(EPH_ERR_FILE, "cannot reread the checkpoint file");
  if ((rc = load_fix(fix, fl, 2))) return rc;
  EPH_HIP(fix, hipStreamSynchronize(fix->stream));
  return EPH_OK;
}

}  // extern "C"
