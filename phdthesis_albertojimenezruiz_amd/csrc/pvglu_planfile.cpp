// pvglu_planfile.cpp -- on-disk form of a symbolic plan.
//
// The pattern half of PV/factorize_electrical_system.h:3-61
// (struct_CPU_CUDA_arrays: filled CSR/CSC, the value maps, the diagonal
// pointers, the level/chain arrays and the per-level widths), plus the
// ordering (P, Prow) it was built with, so a plan can be reused across runs
// and exchanged as a fixture.
//
// Layout, all integers little-endian int32 unless stated:
//   "PVGLUPLN" (8 bytes), version (=2), n, nnz, section count
//   per section: id, length L, then L int32 values
//     ids 0..PVGLU_ARR_COUNT-1: the pvglu_plan_export arrays
//     id 100: P (length n or 0), id 101: Prow (length n or 0)
//   trailing uint64 FNV-1a of every byte before it.
// Loading rebuilds the plan from the filled pattern (sections 0 and 1) and
// requires every other stored array to equal the rebuilt one, so a file
// produced by another build of the same algorithm is checked, not trusted.
#include "pvglu_internal.h"

#include <cstdio>
#include <cstring>
#include <vector>

using namespace pvglu;

namespace {

constexpr char kMagic[8] = {'P', 'V', 'G', 'L', 'U', 'P', 'L', 'N'};
constexpr int kVersion = 2;
constexpr int kSecP = 100, kSecProw = 101;

uint64_t fnv1a(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; i++) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

struct Writer {
  std::vector<unsigned char> buf;
  void put(const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    buf.insert(buf.end(), c, c + n);
  }
  void i32(int v) { put(&v, 4); }
  void section(int id, const int *data, int len) {
    i32(id);
    i32(len);
    if (len > 0)
      put(data, sizeof(int) * (size_t)len);
  }
};

struct Reader {
  const std::vector<unsigned char> &buf;
  size_t pos = 0;
  bool ok = true;
  explicit Reader(const std::vector<unsigned char> &b) : buf(b) {}
  bool get(void *p, size_t n) {
    if (!ok || pos + n > buf.size())
      return ok = false;
    memcpy(p, buf.data() + pos, n);
    pos += n;
    return true;
  }
  int i32() {
    int v = 0;
    get(&v, 4);
    return v;
  }
};

} // namespace

extern "C" int pvglu_plan_save(const pvglu_plan *plan, const char *path,
                               const int *P, const int *Prow) {
  clear_error();
  if (!plan || !path)
    return fail(PVGLU_E_ARG, "plan_save: bad arguments");
  Writer w;
  w.put(kMagic, 8);
  w.i32(kVersion);
  w.i32(plan->n);
  w.i32(plan->nnz);
  w.i32(PVGLU_ARR_COUNT + 2);
  for (int id = 0; id < PVGLU_ARR_COUNT; id++) {
    const int len = pvglu_plan_array_len(plan, id);
    if (len < 0)
      return len;
    std::vector<int> a((size_t)len);
    int st = pvglu_plan_export(plan, id, a.data());
    if (st)
      return st;
    w.section(id, a.data(), len);
  }
  w.section(kSecP, P, P ? plan->n : 0);
  w.section(kSecProw, Prow, Prow ? plan->n : 0);
  const uint64_t h = fnv1a(1469598103934665603ull, w.buf.data(), w.buf.size());
  w.put(&h, 8);
  FILE *f = fopen(path, "wb");
  if (!f)
    return fail(PVGLU_E_ARG, "plan_save: cannot open %s", path);
  const size_t wrote = fwrite(w.buf.data(), 1, w.buf.size(), f);
  const int closed = fclose(f);
  if (wrote != w.buf.size() || closed != 0)
    return fail(PVGLU_E_ARG, "plan_save: short write to %s", path);
  return PVGLU_OK;
}

extern "C" int pvglu_plan_file_n(const char *path) {
  clear_error();
  if (!path)
    return fail(PVGLU_E_ARG, "plan_file_n: bad arguments");
  FILE *f = fopen(path, "rb");
  if (!f)
    return fail(PVGLU_E_ARG, "plan_file_n: cannot open %s", path);
  unsigned char head[16];
  const size_t got = fread(head, 1, sizeof head, f);
  fclose(f);
  int version = 0, n = 0;
  if (got == sizeof head) {
    memcpy(&version, head + 8, 4);
    memcpy(&n, head + 12, 4);
  }
  if (got != sizeof head || memcmp(head, kMagic, 8) != 0 ||
      version != kVersion || n <= 0)
    return fail(PVGLU_E_ARG, "plan_file_n: %s is not a plan file", path);
  return n;
}

extern "C" int pvglu_plan_load(const char *path, pvglu_plan **out, int *P,
                               int *Prow) {
  clear_error();
  if (!path || !out)
    return fail(PVGLU_E_ARG, "plan_load: bad arguments");
  *out = nullptr;
  FILE *f = fopen(path, "rb");
  if (!f)
    return fail(PVGLU_E_ARG, "plan_load: cannot open %s", path);
  std::vector<unsigned char> buf;
  unsigned char tmp[65536];
  size_t got;
  while ((got = fread(tmp, 1, sizeof tmp, f)) > 0)
    buf.insert(buf.end(), tmp, tmp + got);
  fclose(f);
  if (buf.size() < 8 + 16 + 8)
    return fail(PVGLU_E_ARG, "plan_load: %s is too short", path);
  uint64_t stored;
  memcpy(&stored, buf.data() + buf.size() - 8, 8);
  if (fnv1a(1469598103934665603ull, buf.data(), buf.size() - 8) != stored)
    return fail(PVGLU_E_ARG, "plan_load: checksum mismatch in %s", path);
  buf.resize(buf.size() - 8);
  Reader r(buf);
  char magic[8];
  r.get(magic, 8);
  if (!r.ok || memcmp(magic, kMagic, 8) != 0)
    return fail(PVGLU_E_ARG, "plan_load: %s is not a plan file", path);
  const int version = r.i32(), n = r.i32(), nnz = r.i32(), nsec = r.i32();
  if (!r.ok || version != kVersion || n <= 0 || nnz < n || nsec < 2 ||
      nsec > 1024)
    return fail(PVGLU_E_ARG, "plan_load: bad header in %s", path);
  std::vector<std::vector<int>> arr(PVGLU_ARR_COUNT);
  std::vector<int> fileP, fileProw;
  for (int k = 0; k < nsec; k++) {
    const int id = r.i32(), len = r.i32();
    if (!r.ok || len < 0 || (size_t)len * 4 > buf.size())
      return fail(PVGLU_E_ARG, "plan_load: bad section %d in %s", k, path);
    std::vector<int> v((size_t)len);
    if (len > 0 && !r.get(v.data(), sizeof(int) * (size_t)len))
      return fail(PVGLU_E_ARG, "plan_load: truncated section %d", k);
    if (id >= 0 && id < PVGLU_ARR_COUNT)
      arr[id] = std::move(v);
    else if (id == kSecP)
      fileP = std::move(v);
    else if (id == kSecProw)
      fileProw = std::move(v);
  }
  if (r.pos != buf.size())
    return fail(PVGLU_E_ARG, "plan_load: trailing bytes in %s", path);
  const auto &rp = arr[PVGLU_ARR_CSR_ROWPTR], &ci = arr[PVGLU_ARR_CSR_COLIDX];
  if ((int)rp.size() != n + 1 || (int)ci.size() != nnz || rp[n] != nnz)
    return fail(PVGLU_E_ARG, "plan_load: pattern sections do not match n, nnz");
  if ((!fileP.empty() && (int)fileP.size() != n) ||
      (!fileProw.empty() && (int)fileProw.size() != n))
    return fail(PVGLU_E_ARG, "plan_load: ordering sections of wrong length");
  // The checksum only catches accidental corruption: the ordering's values
  // are checked too.  P must be a permutation of [0, n); Prow a permutation
  // that is its own inverse, as the Newton loop's exchange_rows(Prow) only
  // undoes the recorded row swaps then (PV/000_PVPP_simulator.cpp:195-196).
  auto is_perm = [n](const std::vector<int> &p) {
    std::vector<char> seen((size_t)n, 0);
    for (int v : p) {
      if (v < 0 || v >= n || seen[(size_t)v])
        return false;
      seen[(size_t)v] = 1;
    }
    return true;
  };
  if (!fileP.empty() && !is_perm(fileP))
    return fail(PVGLU_E_ARG, "plan_load: stored P of %s is not a permutation",
                path);
  if (!fileProw.empty()) {
    if (!is_perm(fileProw))
      return fail(PVGLU_E_ARG,
                  "plan_load: stored Prow of %s is not a permutation", path);
    for (int i = 0; i < n; i++)
      if (fileProw[(size_t)fileProw[(size_t)i]] != i)
        return fail(PVGLU_E_ARG,
                    "plan_load: stored Prow of %s is not an involution", path);
  }
  pvglu_plan *plan = nullptr;
  int st = pvglu_plan_create(n, rp.data(), ci.data(), &plan);
  if (st)
    return st;
  for (int id = 2; id < PVGLU_ARR_COUNT; id++) {
    const int len = pvglu_plan_array_len(plan, id);
    std::vector<int> mine((size_t)std::max(len, 0));
    if (len < 0 || pvglu_plan_export(plan, id, mine.data()) ||
        mine != arr[id]) {
      pvglu_plan_destroy(plan);
      return fail(PVGLU_E_ARG,
                  "plan_load: array %d of %s differs from the plan rebuilt "
                  "from its pattern",
                  id, path);
    }
  }
  if (P)
    for (int i = 0; i < n; i++)
      P[i] = fileP.empty() ? i : fileP[i];
  if (Prow)
    for (int i = 0; i < n; i++)
      Prow[i] = fileProw.empty() ? i : fileProw[i];
  *out = plan;
  return PVGLU_OK;
}
