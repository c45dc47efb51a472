// Binary files of QuTree's tree containers, byte for byte (host only, no
// device work): TensorTree ("TTre"; TensorTree_Implementation.h:79-132 with
// Tensor::write/read, Tensor_Implementation.h:204-258, and
// TensorShape::write/readDim, TensorShape.cpp:68-111), MatrixTree ("MTre";
// MatrixTree.cpp:52-90) and SparseMatrixTree ("SMTr";
// SparseMatrixTree.cpp:27-70) with Matrix::write/read
// (Matrix_Implementation.h:365-416).  complex<double> coefficients,
// little-endian, as the reference writes them on x86-64.
#include <cstring>
#include <exception>
#include <string>

#include "../../include/qutree_amd.h"
#include "common.hpp"
#include "model.hpp"

namespace qtg {
namespace {

constexpr int32_t kCoeffBytes = 16;  // sizeof(std::complex<double>)

// Appends to buf while there is room; always counts the bytes.
struct Writer {
  unsigned char* buf;
  int64_t cap, n = 0;
  void put(const void* p, size_t len) {
    if (buf && n + (int64_t)len <= cap) std::memcpy(buf + n, p, len);
    n += (int64_t)len;
  }
  void i32(int32_t v) { put(&v, 4); }
  void u64(uint64_t v) { put(&v, 8); }
};

struct Reader {
  const unsigned char* buf;
  int64_t len, n = 0;
  bool get(void* p, size_t k) {
    if (n + (int64_t)k > len) return false;
    std::memcpy(p, buf + n, k);
    n += (int64_t)k;
    return true;
  }
  bool magic(const char* key) {
    char c[4];
    return get(c, 4) && std::memcmp(c, key, 4) == 0;
  }
};

int64_t fail(const std::string& msg, int status) {
  set_error(msg);
  return -(int64_t)status;
}

}  // namespace
}  // namespace qtg

using namespace qtg;

extern "C" {

int64_t qtg_tensor_tree_write(const char* tree_text, const double* psi, unsigned char* buf, int64_t cap) {
  if (!tree_text || !psi || cap < 0) return fail("tensor_tree_write: null argument", QTG_EARG);
  HTree t;
  try {
    t = parse_tree(tree_text);
  } catch (const std::exception& e) {
    return fail(std::string("tensor_tree_write: ") + e.what(), QTG_ESHAPE);
  }
  Writer w{buf, buf ? cap : 0};
  w.put("TTre", 4);
  w.i32((int32_t)t.nodes.size());
  for (const HNode& v : t.nodes) {
    w.put("TENS", 4);
    w.put("TDIM", 4);
    w.i32((int32_t)v.dims.size());
    for (int d : v.dims) w.i32((int32_t)d);
    w.i32(kCoeffBytes);
    w.put(psi + 2 * v.off, (size_t)kCoeffBytes * (size_t)v.N);
  }
  if (buf && w.n > cap) return fail("tensor_tree_write: buffer smaller than " + std::to_string(w.n) + " bytes",
                                    QTG_EARG);
  return w.n;
}

int64_t qtg_tensor_tree_read(const char* tree_text, const unsigned char* buf, int64_t len, double* psi) {
  if (!tree_text || !buf || !psi || len < 0) return fail("tensor_tree_read: null argument", QTG_EARG);
  HTree t;
  try {
    t = parse_tree(tree_text);
  } catch (const std::exception& e) {
    return fail(std::string("tensor_tree_read: ") + e.what(), QTG_ESHAPE);
  }
  Reader r{buf, len};
  if (!r.magic("TTre")) return fail("tensor_tree_read: keyword TTre expected", QTG_ESHAPE);
  int32_t nn = 0;
  if (!r.get(&nn, 4)) return fail("tensor_tree_read: truncated", QTG_ESHAPE);
  if (nn != (int32_t)t.nodes.size())
    return fail("tensor_tree_read: " + std::to_string(nn) + " tensors for a tree of " +
                    std::to_string(t.nodes.size()) + " nodes",
                QTG_ESHAPE);
  for (const HNode& v : t.nodes) {
    if (!r.magic("TENS") || !r.magic("TDIM"))
      return fail("tensor_tree_read: keywords TENS/TDIM expected at node " + std::to_string(v.addr), QTG_ESHAPE);
    int32_t order = 0;
    if (!r.get(&order, 4)) return fail("tensor_tree_read: truncated", QTG_ESHAPE);
    if (order != v.order())
      return fail("tensor_tree_read: node " + std::to_string(v.addr) + " has order " + std::to_string(order) +
                      ", the tree " + std::to_string(v.order()),
                  QTG_ESHAPE);
    for (int k = 0; k < order; ++k) {
      int32_t d = 0;
      if (!r.get(&d, 4)) return fail("tensor_tree_read: truncated", QTG_ESHAPE);
      if (d != v.dims[k])
        return fail("tensor_tree_read: node " + std::to_string(v.addr) + " dimension " + std::to_string(k) +
                        " is " + std::to_string(d) + ", the tree's " + std::to_string(v.dims[k]),
                    QTG_ESHAPE);
    }
    int32_t size = 0;
    if (!r.get(&size, 4)) return fail("tensor_tree_read: truncated", QTG_ESHAPE);
    if (size != kCoeffBytes)
      return fail("tensor_tree_read: coefficients of " + std::to_string(size) + " bytes, complex<double> expected",
                  QTG_ESHAPE);
    if (!r.get(psi + 2 * v.off, (size_t)kCoeffBytes * (size_t)v.N))
      return fail("tensor_tree_read: truncated", QTG_ESHAPE);
  }
  return r.n;
}

int64_t qtg_matrix_tree_write(int sparse, int64_t count, const int64_t* dims, const double* mats, unsigned char* buf,
                              int64_t cap) {
  if (count < 0 || (count > 0 && (!dims || !mats)) || cap < 0)
    return fail("matrix_tree_write: null argument", QTG_EARG);
  Writer w{buf, buf ? cap : 0};
  w.put(sparse ? "SMTr" : "MTre", 4);
  w.i32((int32_t)count);
  int64_t off = 0;
  for (int64_t m = 0; m < count; ++m) {
    const int64_t d1 = dims[2 * m], d2 = dims[2 * m + 1];
    if (d1 < 0 || d2 < 0) return fail("matrix_tree_write: negative dimension", QTG_EARG);
    w.put("MATR", 4);
    w.u64((uint64_t)d1);
    w.u64((uint64_t)d2);
    w.i32(kCoeffBytes);
    w.put(mats + 2 * off, (size_t)kCoeffBytes * (size_t)(d1 * d2));
    off += d1 * d2;
  }
  if (buf && w.n > cap) return fail("matrix_tree_write: buffer smaller than " + std::to_string(w.n) + " bytes",
                                    QTG_EARG);
  return w.n;
}

int64_t qtg_matrix_tree_read(int sparse, const unsigned char* buf, int64_t len, int64_t max_count, int64_t* dims,
                             double* mats, int64_t mats_cap, int64_t* count) {
  if (!buf || len < 0 || !count) return fail("matrix_tree_read: null argument", QTG_EARG);
  Reader r{buf, len};
  if (!r.magic(sparse ? "SMTr" : "MTre"))
    return fail(std::string("matrix_tree_read: keyword ") + (sparse ? "SMTr" : "MTre") + " expected", QTG_ESHAPE);
  int32_t nn = 0;
  if (!r.get(&nn, 4) || nn < 0) return fail("matrix_tree_read: truncated", QTG_ESHAPE);
  *count = nn;
  if (nn > max_count || (nn > 0 && !dims))
    return fail("matrix_tree_read: " + std::to_string(nn) + " matrices, room for " + std::to_string(max_count),
                QTG_EARG);
  int64_t off = 0;
  for (int32_t m = 0; m < nn; ++m) {
    uint64_t d1 = 0, d2 = 0;
    int32_t size = 0;
    if (!r.magic("MATR")) return fail("matrix_tree_read: keyword MATR expected", QTG_ESHAPE);
    if (!r.get(&d1, 8) || !r.get(&d2, 8) || !r.get(&size, 4)) return fail("matrix_tree_read: truncated", QTG_ESHAPE);
    if (size != kCoeffBytes)
      return fail("matrix_tree_read: coefficients of " + std::to_string(size) + " bytes, complex<double> expected",
                  QTG_ESHAPE);
    if (d1 > (1u << 30) || d2 > (1u << 30)) return fail("matrix_tree_read: implausible dimensions", QTG_ESHAPE);
    const int64_t e = (int64_t)(d1 * d2);
    if (off + e > mats_cap || (e > 0 && !mats))
      return fail("matrix_tree_read: more coefficients than mats_cap", QTG_EARG);
    dims[2 * m] = (int64_t)d1;
    dims[2 * m + 1] = (int64_t)d2;
    if (e > 0 && !r.get(mats + 2 * off, (size_t)kCoeffBytes * (size_t)e))
      return fail("matrix_tree_read: truncated", QTG_ESHAPE);
    off += e;
  }
  return r.n;
}

}  // extern "C"
