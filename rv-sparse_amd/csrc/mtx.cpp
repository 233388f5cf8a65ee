// mtx.cpp — Matrix Market reader and CSR canonicalisation (host code).
//
// Same matrix as the reference's bench-tool converter
// (tools/src/mtx_to_csr_formatter.c:74-325) for well-formed files; the
// work is split over host threads: the body is cut at line boundaries and
// parsed in parallel, and rows are sorted and folded in parallel. Entry
// order (which decides the order in which repeated entries are summed) is
// file order throughout: chunks are concatenated in order, the scatter
// into rows is sequential, and the per-row sort is stable.
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <string>
#include <thread>
#include <vector>

#include "rv_sparse_mtx.h"

namespace {

struct Entries {
  std::vector<int32_t> row, col;
  std::vector<float> val;
};

int thread_count(int32_t requested) {
  if (requested > 0) return std::min(requested, 64);
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
}

template <class F>
void parallel_for(int n, int threads, F&& f) {
  if (threads <= 1 || n <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  const int t = std::min(threads, n);
  for (int w = 0; w < t; ++w)
    pool.emplace_back([&, w] {
      for (int i = w; i < n; i += t) f(i);
    });
  for (std::thread& th : pool) th.join();
}

bool read_all(const char* path, std::string& buf) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  buf.clear();
  char tmp[1 << 16];
  size_t got;
  while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.append(tmp, got);
  const bool ok = !ferror(f);
  fclose(f);
  return ok;
}

// Next line [p, e) of buf starting at pos; pos moves past its newline.
bool next_line(const std::string& buf, size_t& pos, size_t& p, size_t& e) {
  if (pos >= buf.size()) return false;
  p = pos;
  const size_t nl = buf.find('\n', pos);
  e = nl == std::string::npos ? buf.size() : nl;
  pos = nl == std::string::npos ? buf.size() : nl + 1;
  return true;
}

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f'; }

// Parses "row col [value]" entries from [p, e) (e at a line boundary or the
// end of the text, which is NUL-terminated). Returns false on a malformed
// token.
bool parse_chunk(const char* p, const char* e, bool pattern, Entries& out) {
  while (true) {
    while (p < e && is_space(*p)) ++p;
    if (p >= e) return true;
    // Tokens may not run past the chunk (a line missing its value is
    // malformed rather than borrowing the next line's first token).
    char* q = nullptr;
    errno = 0;
    const long r = strtol(p, &q, 10);
    if (q == p || q > e || errno) return false;
    p = q;
    const long c = strtol(p, &q, 10);
    if (q == p || q > e || errno) return false;
    p = q;
    float v = 1.0f;
    if (!pattern) {
      v = strtof(p, &q);  // the reference's fscanf("%f"): decimal text to FP32
      if (q == p || q > e) return false;
      p = q;
    }
    if (r < 1 || r > INT_MAX || c < 1 || c > INT_MAX) return false;
    out.row.push_back((int32_t)(r - 1));
    out.col.push_back((int32_t)(c - 1));
    out.val.push_back(v);
  }
}

// Sorts every row of (rp, ci, vv) by column (stable) and folds repeated
// columns, summing their values in order; writes the folded matrix to
// (out_rp, out_ci, out_v), which may not alias the input. Returns nnz.
long long canonicalise(int rows, const int32_t* rp, int32_t* ci, float* vv, int32_t* out_rp,
                       int32_t* out_ci, float* out_v, int threads) {
  std::vector<int32_t> len(rows, 0);
  const int blocks = std::max(1, std::min(rows, threads * 8));
  // Pass 1: sort each row in place (column, value pairs; stable), count the
  // distinct columns.
  parallel_for(blocks, threads, [&](int b) {
    const int r0 = (int)((long long)rows * b / blocks), r1 = (int)((long long)rows * (b + 1) / blocks);
    std::vector<std::pair<int32_t, float>> buf;
    for (int i = r0; i < r1; ++i) {
      const int s = rp[i], n = rp[i + 1] - rp[i];
      buf.resize(n);
      for (int t = 0; t < n; ++t) buf[t] = {ci[s + t], vv[s + t]};
      std::stable_sort(buf.begin(), buf.end(),
                       [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& b) {
                         return a.first < b.first;
                       });
      int distinct = 0;
      for (int t = 0; t < n; ++t) {
        ci[s + t] = buf[t].first;
        vv[s + t] = buf[t].second;
        if (t == 0 || buf[t].first != buf[t - 1].first) ++distinct;
      }
      len[i] = distinct;
    }
  });
  long long total = 0;
  out_rp[0] = 0;
  for (int i = 0; i < rows; ++i) {
    total += len[i];
    out_rp[i + 1] = (int32_t)total;
  }
  // Pass 2: fold runs of one column, first value plus the next, and so on.
  parallel_for(blocks, threads, [&](int b) {
    const int r0 = (int)((long long)rows * b / blocks), r1 = (int)((long long)rows * (b + 1) / blocks);
    for (int i = r0; i < r1; ++i) {
      int w = out_rp[i];
      const int s = rp[i], e = rp[i + 1];
      int t = s;
      while (t < e) {
        const int32_t col = ci[t];
        float acc = vv[t];
        ++t;
        while (t < e && ci[t] == col) {
          acc += vv[t];
          ++t;
        }
        out_ci[w] = col;
        out_v[w] = acc;
        ++w;
      }
    }
  });
  return total;
}

rvsp_status_t read_mtx(const char* path, int threads, rvsp_csr_matrix_t* out) {
  std::string text;
  if (!read_all(path, text)) return RVSP_ERROR_INVALID_ARGUMENT;
  size_t pos = 0, p = 0, e = 0;
  // Banner: %%MatrixMarket object format field symmetry.
  if (!next_line(text, pos, p, e)) return RVSP_ERROR_INVALID_ARGUMENT;
  const std::string banner = text.substr(p, e - p);
  char object[64], format[64], field[64], symmetry[64];
  if (banner.compare(0, 14, "%%MatrixMarket") != 0 ||
      sscanf(banner.c_str(), "%%%%MatrixMarket %63s %63s %63s %63s", object, format, field,
             symmetry) < 4)
    return RVSP_ERROR_INVALID_ARGUMENT;
  if (strcmp(format, "coordinate") != 0 || strcmp(field, "complex") == 0)
    return RVSP_ERROR_INVALID_ARGUMENT;
  const bool pattern = strcmp(field, "pattern") == 0;
  const bool symmetric = strcmp(symmetry, "symmetric") == 0;
  // Comments, then the size line.
  std::string size_line;
  while (next_line(text, pos, p, e)) {
    if (e > p && text[p] == '%') continue;
    size_line = text.substr(p, e - p);
    break;
  }
  long long M = 0, N = 0, L = 0;
  if (sscanf(size_line.c_str(), "%lld %lld %lld", &M, &N, &L) != 3 || M < 1 || N < 1 || L < 0 ||
      M > INT_MAX - 1 || N > INT_MAX || L > INT_MAX)
    return RVSP_ERROR_INVALID_ARGUMENT;
  // Body: chunks at line boundaries, parsed in parallel, kept in order.
  const size_t body = pos, end = text.size();
  const int nchunks = std::max(1, std::min<int>(threads * 4, (int)((end - body) >> 16) + 1));
  std::vector<size_t> cut(nchunks + 1, end);
  cut[0] = body;
  for (int k = 1; k < nchunks; ++k) {
    size_t c = body + (end - body) * k / nchunks;
    c = std::max(c, cut[k - 1]);
    const size_t nl = text.find('\n', c);
    cut[k] = nl == std::string::npos ? end : nl + 1;
  }
  std::vector<Entries> parts(nchunks);
  std::vector<char> ok(nchunks, 1);
  parallel_for(nchunks, threads, [&](int k) {
    ok[k] = parse_chunk(text.data() + cut[k], text.data() + cut[k + 1], pattern, parts[k]);
  });
  long long got = 0;
  for (int k = 0; k < nchunks; ++k) {
    if (!ok[k]) return RVSP_ERROR_INVALID_ARGUMENT;
    got += (long long)parts[k].row.size();
  }
  if (got < L) return RVSP_ERROR_INVALID_ARGUMENT;  // truncated
  // The first L entries in file order, then the mirrors of the
  // off-diagonal ones for symmetric storage.
  Entries all;
  try {
    all.row.reserve(symmetric ? 2 * L : L);
    all.col.reserve(symmetric ? 2 * L : L);
    all.val.reserve(symmetric ? 2 * L : L);
    for (int k = 0; k < nchunks && (long long)all.row.size() < L; ++k) {
      const long long take = std::min<long long>(L - (long long)all.row.size(),
                                                 (long long)parts[k].row.size());
      all.row.insert(all.row.end(), parts[k].row.begin(), parts[k].row.begin() + take);
      all.col.insert(all.col.end(), parts[k].col.begin(), parts[k].col.begin() + take);
      all.val.insert(all.val.end(), parts[k].val.begin(), parts[k].val.begin() + take);
      Entries().row.swap(parts[k].row);
      Entries().col.swap(parts[k].col);
      Entries().val.swap(parts[k].val);
    }
  } catch (...) {
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
  for (long long k = 0; k < L; ++k)
    if (all.row[k] >= M || all.col[k] >= N) return RVSP_ERROR_INVALID_ARGUMENT;
  if (symmetric)
    for (long long k = 0; k < L; ++k)
      if (all.row[k] != all.col[k]) {
        if (all.col[k] >= M || all.row[k] >= N) return RVSP_ERROR_INVALID_ARGUMENT;
        all.row.push_back(all.col[k]);
        all.col.push_back(all.row[k]);
        all.val.push_back(all.val[k]);
      }
  const long long nnz = (long long)all.row.size();
  if (nnz > INT_MAX) return RVSP_ERROR_INVALID_ARGUMENT;
  // Rows by counting; the scatter keeps file order within every row.
  std::vector<int32_t> rp(M + 1, 0), ci(nnz), cur;
  std::vector<float> vv(nnz);
  for (long long k = 0; k < nnz; ++k) ++rp[all.row[k] + 1];
  for (long long i = 0; i < M; ++i) rp[i + 1] += rp[i];
  cur.assign(rp.begin(), rp.end() - 1);
  for (long long k = 0; k < nnz; ++k) {
    const int32_t d = cur[all.row[k]]++;
    ci[d] = all.col[k];
    vv[d] = all.val[k];
  }
  Entries().row.swap(all.row);
  Entries().col.swap(all.col);
  Entries().val.swap(all.val);
  int32_t* o_rp = (int32_t*)malloc(sizeof(int32_t) * (M + 1));
  int32_t* o_ci = (int32_t*)malloc(sizeof(int32_t) * std::max<long long>(nnz, 1));
  float* o_v = (float*)malloc(sizeof(float) * std::max<long long>(nnz, 1));
  if (!o_rp || !o_ci || !o_v) {
    free(o_rp);
    free(o_ci);
    free(o_v);
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
  const long long folded = canonicalise((int)M, rp.data(), ci.data(), vv.data(), o_rp, o_ci,
                                        o_v, threads);
  memset(out, 0, sizeof *out);
  out->rows = (int32_t)M;
  out->cols = (int32_t)N;
  out->nnz = (int32_t)folded;
  out->row_ptr = o_rp;
  out->col_idx = o_ci;
  out->values = o_v;
  out->dtype = RVSP_DTYPE_FP32;
  out->format = RVSP_FORMAT_CSR;
  out->owns_data = 1;
  return RVSP_SUCCESS;
}

}  // namespace

extern "C" {

rvsp_status_t rvsp_csr_read_mtx(const char* path, int32_t threads, rvsp_csr_matrix_t* out) {
  if (path == NULL || out == NULL) return RVSP_ERROR_NULL_POINTER;
  memset(out, 0, sizeof *out);
  try {
    return read_mtx(path, thread_count(threads), out);
  } catch (const std::bad_alloc&) {
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
}

rvsp_status_t rvsp_csr_canonicalize(rvsp_csr_matrix_t* A) {
  if (A == NULL || A->row_ptr == NULL || A->col_idx == NULL || A->values == NULL)
    return RVSP_ERROR_NULL_POINTER;
  if (A->dtype != RVSP_DTYPE_FP32) return RVSP_ERROR_UNSUPPORTED_DTYPE;
  const int rows = A->rows;
  if (rows < 0 || A->cols < 0 || A->row_ptr[0] != 0) return RVSP_ERROR_INVALID_ARGUMENT;
  for (int i = 0; i < rows; ++i)
    if (A->row_ptr[i + 1] < A->row_ptr[i]) return RVSP_ERROR_INVALID_ARGUMENT;
  const int32_t n = A->row_ptr[rows];
  if (n > A->nnz && A->nnz >= 0) return RVSP_ERROR_INVALID_ARGUMENT;
  for (int32_t k = 0; k < n; ++k)
    if (A->col_idx[k] < 0 || A->col_idx[k] >= A->cols) return RVSP_ERROR_INVALID_ARGUMENT;
  try {
    std::vector<int32_t> rp(A->row_ptr, A->row_ptr + rows + 1), ci(n);
    std::vector<float> vv(n);
    const int threads = thread_count(0);
    const long long folded = canonicalise(rows, rp.data(), A->col_idx, (float*)A->values,
                                          A->row_ptr, ci.data(), vv.data(), threads);
    memcpy(A->col_idx, ci.data(), sizeof(int32_t) * folded);
    memcpy(A->values, vv.data(), sizeof(float) * folded);
    A->nnz = (int32_t)folded;
  } catch (const std::bad_alloc&) {
    return RVSP_ERROR_ALLOCATION_FAILED;
  }
  return RVSP_SUCCESS;
}

}  // extern "C"
