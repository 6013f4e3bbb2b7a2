// checkpoint.hip -- particle backup files in the reference's format.
//
// The reference writes DATA.BAK/particlesNNN.bak (main.f90:1102-1132) and
// reads it back in restart (restart.f90:56-64, and 66-113 for files written
// with another number of OpenMP threads) as Fortran unformatted sequential
// records:
//   record 1: time (real*8), gtype (integer*4), Pabs_cor (real*8)
//   record 2: np_tot(1:ntype,1:nproc) (integer*4, ntype fastest)
//   then for iproc = 1..nproc, ptype = 1..ntype: vxp(1:6,1:np_tot(ptype,iproc))
// Each record is framed by 4-byte length markers (the gfortran/flang
// convention).  A record longer than 2^31-9 bytes is split into subrecords:
// a negative header says that another subrecord follows, a negative footer
// that one came before.
//
// Host code only: the context entry points move particles through
// lepic_download_particles / lepic_upload_particles.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lepic_internal.h"

namespace {

constexpr uint64_t kMaxSub = 2147483639ull;   // largest subrecord payload

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() { if (f) std::fclose(f); }
};

bool put_i32(FILE* f, int32_t v) { return std::fwrite(&v, 4, 1, f) == 1; }
bool get_i32(FILE* f, int32_t* v) { return std::fread(v, 4, 1, f) == 1; }

bool write_record(FILE* f, const void* data, uint64_t bytes) {
  const char* p = (const char*)data;
  uint64_t left = bytes;
  bool first = true;
  do {
    const uint64_t n = left > kMaxSub ? kMaxSub : left;
    const bool last = n == left;
    if (!put_i32(f, last ? (int32_t)n : -(int32_t)n)) return false;
    if (n && std::fwrite(p, 1, n, f) != n) return false;
    if (!put_i32(f, first ? (int32_t)n : -(int32_t)n)) return false;
    p += n;
    left -= n;
    first = false;
  } while (left > 0);
  return true;
}

// Reads one record into out (resized); false on a malformed or short file.
bool read_record(FILE* f, std::vector<char>& out) {
  out.clear();
  for (;;) {
    int32_t head = 0, foot = 0;
    if (!get_i32(f, &head)) return false;
    const uint64_t n = (uint64_t)(head < 0 ? -(int64_t)head : head);
    const size_t at = out.size();
    out.resize(at + n);
    if (n && std::fread(out.data() + at, 1, n, f) != n) return false;
    if (!get_i32(f, &foot)) return false;
    if ((uint64_t)(foot < 0 ? -(int64_t)foot : foot) != n) return false;
    if (head >= 0) return true;
  }
}

// Skips one record (possibly split); returns its payload size or -1.
int64_t skip_record(FILE* f) {
  int64_t total = 0;
  for (;;) {
    int32_t head = 0, foot = 0;
    if (!get_i32(f, &head)) return -1;
    const int64_t n = head < 0 ? -(int64_t)head : head;
    if (std::fseek(f, (long)n, SEEK_CUR) != 0) return -1;
    if (!get_i32(f, &foot)) return -1;
    if ((foot < 0 ? -(int64_t)foot : foot) != n) return -1;
    total += n;
    if (head >= 0) return total;
  }
}

// Header records; nproc is the file's number of partitions (record 2 holds
// ntype*nproc counts).
int read_header(FILE* f, int ntype, double* time, int32_t* gtype, double* pabs, std::vector<int32_t>& np_tot) {
  std::vector<char> rec;
  if (!read_record(f, rec) || rec.size() != 20) return LEPIC_ERR_ARG;
  if (time) std::memcpy(time, rec.data(), 8);
  if (gtype) std::memcpy(gtype, rec.data() + 8, 4);
  if (pabs) std::memcpy(pabs, rec.data() + 12, 8);
  if (!read_record(f, rec) || rec.size() % (4 * (size_t)ntype) != 0 || rec.empty()) return LEPIC_ERR_ARG;
  np_tot.resize(rec.size() / 4);
  std::memcpy(np_tot.data(), rec.data(), rec.size());
  for (int32_t v : np_tot)
    if (v < 0) return LEPIC_ERR_ARG;
  return 0;
}

// Appends the particles of ptype from every partition of the file, in
// partition order (the flag_restart=2 path of restart.f90:66-113 with one
// thread), to out.
int read_species(FILE* f, int ntype, int ptype, std::vector<double>& out) {
  std::vector<int32_t> np_tot;
  if (int rc = read_header(f, ntype, nullptr, nullptr, nullptr, np_tot)) return rc;
  const int nproc = (int)np_tot.size() / ntype;
  std::vector<char> rec;
  for (int ip = 0; ip < nproc; ++ip)
    for (int p = 1; p <= ntype; ++p) {
      const int64_t n = np_tot[(size_t)ip * ntype + (p - 1)];
      if (p != ptype) {
        if (skip_record(f) != n * 48) return LEPIC_ERR_ARG;
        continue;
      }
      if (!read_record(f, rec) || (int64_t)rec.size() != n * 48) return LEPIC_ERR_ARG;
      const size_t at = out.size();
      out.resize(at + (size_t)n * 6);
      if (n) std::memcpy(out.data() + at, rec.data(), rec.size());
    }
  return 0;
}

}  // namespace

extern "C" {

int lepic_backup_write(const char* path, double time, int32_t gtype, double pabs_cor, int32_t ntype, int32_t nproc,
                       const int32_t* np_tot, const double* const* vxp) {
  if (!path || ntype < 1 || nproc < 1 || !np_tot || !vxp) return LEPIC_ERR_ARG;
  for (int64_t k = 0; k < (int64_t)ntype * nproc; ++k)
    if (np_tot[k] < 0 || (np_tot[k] > 0 && !vxp[k])) return LEPIC_ERR_ARG;
  File F(path, "wb");
  if (!F.f) return LEPIC_ERR_ARG;
  char head[20];
  std::memcpy(head, &time, 8);
  std::memcpy(head + 8, &gtype, 4);
  std::memcpy(head + 12, &pabs_cor, 8);
  if (!write_record(F.f, head, 20)) return LEPIC_ERR_ARG;
  if (!write_record(F.f, np_tot, 4 * (uint64_t)ntype * nproc)) return LEPIC_ERR_ARG;
  for (int ip = 0; ip < nproc; ++ip)
    for (int p = 0; p < ntype; ++p) {
      const int64_t k = (int64_t)ip * ntype + p;
      if (!write_record(F.f, vxp[k], 48 * (uint64_t)np_tot[k])) return LEPIC_ERR_ARG;
    }
  return std::fflush(F.f) == 0 ? 0 : LEPIC_ERR_ARG;
}

int lepic_backup_header(const char* path, int32_t ntype, double* time, int32_t* gtype, double* pabs_cor,
                        int32_t* nproc, int32_t* np_tot, int32_t cap) {
  if (!path || ntype < 1 || !nproc) return LEPIC_ERR_ARG;
  File F(path, "rb");
  if (!F.f) return LEPIC_ERR_ARG;
  std::vector<int32_t> counts;
  if (int rc = read_header(F.f, ntype, time, gtype, pabs_cor, counts)) return rc;
  *nproc = (int32_t)counts.size() / ntype;
  if (np_tot) {
    if ((int32_t)counts.size() > cap) return LEPIC_ERR_CAPACITY;
    std::memcpy(np_tot, counts.data(), counts.size() * 4);
  }
  return 0;
}

int lepic_backup_read_species(const char* path, int32_t ntype, int32_t ptype, double* vxp, int64_t cap,
                              int64_t* np) {
  if (!path || ntype < 1 || ptype < 1 || ptype > ntype || !np) return LEPIC_ERR_ARG;
  File F(path, "rb");
  if (!F.f) return LEPIC_ERR_ARG;
  std::vector<double> all;
  if (int rc = read_species(F.f, ntype, ptype, all)) return rc;
  *np = (int64_t)(all.size() / 6);
  if (*np > cap) return LEPIC_ERR_CAPACITY;
  if (*np && !vxp) return LEPIC_ERR_ARG;
  if (*np) std::memcpy(vxp, all.data(), all.size() * 8);
  return 0;
}

int lepic_save_particles(lepic_ctx* ctx, const char* path, double time, int32_t gtype, double pabs_cor) {
  if (!ctx || !path) return LEPIC_ERR_ARG;
  const int ntype = ctx->nsp;
  std::vector<std::vector<double>> buf(ntype);
  std::vector<int32_t> np_tot(ntype);
  std::vector<const double*> ptr(ntype);
  for (int p = 1; p <= ntype; ++p) {
    int64_t n = 0;
    if (int rc = lepic_get_count(ctx, p, &n)) return rc;
    if (n > 2147483647) return ctx->fail(LEPIC_ERR_CAPACITY, "np_tot exceeds the file format's integer*4");
    buf[p - 1].resize((size_t)n * 6);
    int64_t got = 0;
    if (int rc = lepic_download_particles(ctx, p, buf[p - 1].data(), n, &got)) return rc;
    np_tot[p - 1] = (int32_t)got;
    ptr[p - 1] = buf[p - 1].data();
  }
  const int rc = lepic_backup_write(path, time, gtype, pabs_cor, ntype, 1, np_tot.data(), ptr.data());
  if (rc) return ctx->fail(rc, std::string("cannot write backup file ") + path);
  return 0;
}

int lepic_load_particles(lepic_ctx* ctx, const char* path, double* time, int32_t* gtype, double* pabs_cor) {
  if (!ctx || !path) return LEPIC_ERR_ARG;
  const int ntype = ctx->nsp;
  int32_t nproc = 0;
  if (lepic_backup_header(path, ntype, time, gtype, pabs_cor, &nproc, nullptr, 0))
    return ctx->fail(LEPIC_ERR_ARG, std::string("backup files were not read correctly: ") + path);
  for (int p = 1; p <= ntype; ++p) {
    std::vector<double> v;
    {
      File F(path, "rb");
      if (!F.f || read_species(F.f, ntype, p, v))
        return ctx->fail(LEPIC_ERR_ARG, std::string("backup files were not read correctly: ") + path);
    }
    const int64_t n = (int64_t)(v.size() / 6);
    if (n > ctx->sp[p - 1].cap) return ctx->fail(LEPIC_ERR_CAPACITY, "np_tot > nmax in restart()");
    if (int rc = lepic_upload_particles(ctx, p, v.data(), n)) return rc;
  }
  return 0;
}

}  // extern "C"
