// bindings.cpp -- pybind11 module turboquant._turboquant, the drop-in for the
// reference's python/turboquant/bindings.cpp. Same class name, constructor
// arguments and defaults, method names, argument names, return shapes and
// ValueError messages; the work goes through the C ABI of libtq_hip.so
// (include/tq_hip.h) to the gfx950 kernels. The GIL is released while the
// device works.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tq_hip.h"

namespace py = pybind11;

namespace {

void check(int rc) {
  if (rc == TQ_OK) return;
  if (rc == TQ_EINVAL) throw py::value_error(tq_last_error());
  throw std::runtime_error(tq_last_error());
}

// Validation of float inputs as in bindings.cpp:14-36; rows must also be
// contiguous (the reference accepted row-strided views and misread them).
const float *as_float_ptr(const py::buffer_info &info, ssize_t expected_last,
                          const char *name, int expected_ndim, ssize_t *out_rows = nullptr) {
  if (info.format != py::format_descriptor<float>::format())
    throw py::value_error(std::string(name) + ": expected float32");
  if (info.ndim != expected_ndim) throw py::value_error(std::string(name) + ": wrong ndim");
  if (expected_ndim == 1) {
    if (info.shape[0] != expected_last)
      throw py::value_error(std::string(name) + ": wrong length");
  } else {
    if (info.shape[1] != expected_last) throw py::value_error(std::string(name) + ": wrong dim");
    if (out_rows) *out_rows = info.shape[0];
  }
  ssize_t stride_last = info.strides[info.ndim - 1];
  bool rows_ok = info.ndim == 1 || info.shape[0] <= 1 ||
                 info.strides[0] == static_cast<ssize_t>(sizeof(float)) * expected_last;
  if (stride_last != static_cast<ssize_t>(sizeof(float)) || !rows_ok)
    throw py::value_error(std::string(name) + ": must be C-contiguous float32");
  return static_cast<const float *>(info.ptr);
}

// bindings.cpp:38-49
void *as_bytes_ptr(const py::buffer_info &info, ssize_t expected_bytes, const char *name) {
  if (info.itemsize != 1) throw py::value_error(std::string(name) + ": must be uint8");
  ssize_t total = 1;
  for (auto s : info.shape) total *= s;
  if (total != expected_bytes) throw py::value_error(std::string(name) + ": wrong byte size");
  return info.ptr;
}

ssize_t total_of(const py::buffer_info &info) {
  ssize_t total = 1;
  for (auto s : info.shape) total *= s;
  return total;
}

int parse_mode(const std::string &mode) {
  if (mode == "fp32") return TQ_MODE_FP32;
  if (mode == "mfma") return TQ_MODE_MFMA;
  throw py::value_error("mode must be 'fp32' or 'mfma'");
}

// Opaque prepared forms (bindings.cpp:56-57): host copies of the rows the
// C ABI's prepare calls write.
struct PreparedQuery {
  std::vector<float> row;  // [q_rot | s_q | q_norm_sq | q_norm | 0 | 0]
};
struct PreparedSymCode {
  std::vector<float> row;  // tq_prepare_symmetric's row
};

class Space {
 public:
  Space(size_t dim, int bits_per_coord, uint64_t rot_seed, uint64_t qjl_seed, int num_threads)
      : num_threads_(num_threads > 0 ? num_threads
                                     : static_cast<int>(std::thread::hardware_concurrency())) {
    check(tq_space_create(dim, bits_per_coord, rot_seed, qjl_seed, &h_));
    check(tq_space_info_get(h_, &info_));
    if (info_.padded)
      std::cerr << "[TurboQuantSpace] warning: input dim " << info_.dim
                << " is not a power of 2 (or < 4); padding to " << info_.padded_dim
                << " with zeros for internal computation.\n";
  }
  ~Space() { tq_space_destroy(h_); }
  Space(const Space &) = delete;
  Space &operator=(const Space &) = delete;

  size_t dim() const { return info_.dim; }
  size_t padded_dim() const { return info_.padded_dim; }
  bool padded() const { return info_.padded != 0; }
  int num_threads() const { return num_threads_; }
  size_t code_size() const { return info_.code_size; }
  int bits() const { return info_.bits_per_coord; }

  py::array_t<uint8_t> encode(py::buffer x) {
    auto info = x.request();
    const float *p = as_float_ptr(info, dim(), "input", 1);
    py::array_t<uint8_t> code(static_cast<ssize_t>(code_size()));
    uint8_t *o = code.mutable_data();
    run([&] { return tq_encode_batch(h_, p, 1, o, TQ_MEM_HOST, nullptr); });
    return code;
  }

  void encode_into(py::buffer x, py::buffer out) {
    auto info = x.request();
    const float *p = as_float_ptr(info, dim(), "input", 1);
    auto oinfo = out.request(true);
    auto *o = static_cast<uint8_t *>(as_bytes_ptr(oinfo, code_size(), "code"));
    run([&] { return tq_encode_batch(h_, p, 1, o, TQ_MEM_HOST, nullptr); });
  }

  py::array_t<uint8_t> encode_batch(py::buffer X) {
    auto info = X.request();
    ssize_t rows = 0;
    const float *p = as_float_ptr(info, dim(), "input", 2, &rows);
    py::array_t<uint8_t> codes({rows, static_cast<ssize_t>(code_size())});
    uint8_t *o = codes.mutable_data();
    run([&] { return tq_encode_batch(h_, p, (uint64_t)rows, o, TQ_MEM_HOST, nullptr); });
    return codes;
  }

  void encode_batch_into(py::buffer X, py::buffer out) {
    auto info = X.request();
    ssize_t rows = 0;
    const float *p = as_float_ptr(info, dim(), "input", 2, &rows);
    auto oinfo = out.request(true);
    auto *o = static_cast<uint8_t *>(
        as_bytes_ptr(oinfo, rows * static_cast<ssize_t>(code_size()), "codes"));
    run([&] { return tq_encode_batch(h_, p, (uint64_t)rows, o, TQ_MEM_HOST, nullptr); });
  }

  float distance(py::buffer query, py::buffer code) {
    auto qi = query.request();
    const float *q = as_float_ptr(qi, dim(), "query", 1);
    auto ci = code.request();
    const auto *c = static_cast<const uint8_t *>(as_bytes_ptr(ci, code_size(), "code"));
    float d = 0.0f;
    run([&] { return tq_distance_1_to_n(h_, q, c, 1, &d, TQ_MEM_HOST, nullptr); });
    return d;
  }

  float distance_symmetric(py::buffer a, py::buffer b) {
    auto ai = a.request();
    auto bi = b.request();
    const auto *ca = static_cast<const uint8_t *>(as_bytes_ptr(ai, code_size(), "code_a"));
    const auto *cb = static_cast<const uint8_t *>(as_bytes_ptr(bi, code_size(), "code_b"));
    float d = 0.0f;
    run([&] { return tq_distance_m_to_n_symmetric(h_, ca, 1, cb, 1, &d, TQ_MEM_HOST, nullptr); });
    return d;
  }

  py::array_t<float> distance_1_to_n(py::buffer query, py::buffer codes) {
    auto qi = query.request();
    const float *q = as_float_ptr(qi, dim(), "query", 1);
    auto ci = codes.request();
    const size_t n = code_count(ci);
    py::array_t<float> out(static_cast<ssize_t>(n));
    float *o = out.mutable_data();
    const auto *c = static_cast<const uint8_t *>(ci.ptr);
    run([&] { return tq_distance_1_to_n(h_, q, c, n, o, TQ_MEM_HOST, nullptr); });
    return out;
  }

  py::array_t<float> distance_m_to_n(py::buffer queries, py::buffer codes,
                                     const std::string &mode) {
    auto qi = queries.request();
    ssize_t m = 0;
    const float *q = as_float_ptr(qi, dim(), "queries", 2, &m);
    auto ci = codes.request();
    const size_t n = code_count(ci);
    const int md = parse_mode(mode);
    py::array_t<float> out({m, static_cast<ssize_t>(n)});
    float *o = out.mutable_data();
    const auto *c = static_cast<const uint8_t *>(ci.ptr);
    run([&] { return tq_distance_m_to_n(h_, q, (uint64_t)m, c, n, o, md, TQ_MEM_HOST, nullptr); });
    return out;
  }

  py::array_t<float> distance_m_to_n_symmetric(py::buffer codes_a, py::buffer codes_b,
                                               const std::string &mode) {
    return sym_matrix(codes_a, codes_b, TQ_SYM_MSE, parse_mode(mode));
  }
  // bindings.cpp:221-266
  py::array_t<float> distance_m_to_n_symmetric_full(py::buffer codes_a, py::buffer codes_b) {
    return sym_matrix(codes_a, codes_b, TQ_SYM_FULL, TQ_MODE_FP32);
  }
  py::array_t<float> distance_m_to_n_symmetric_light(py::buffer codes_a, py::buffer codes_b) {
    return sym_matrix(codes_a, codes_b, TQ_SYM_LIGHT, TQ_MODE_FP32);
  }

  // bindings.cpp:119-132
  py::array_t<float> rotated_coords_batch(py::buffer X) {
    auto info = X.request();
    ssize_t rows = 0;
    const float *p = as_float_ptr(info, dim(), "input", 2, &rows);
    py::array_t<float> out({rows, static_cast<ssize_t>(padded_dim())});
    float *o = out.mutable_data();
    run([&] { return tq_rotated_coords_batch(h_, p, (uint64_t)rows, o, TQ_MEM_HOST, nullptr); });
    return out;
  }

  // bindings.cpp:268-274
  PreparedQuery prepare_query(py::buffer query) {
    auto qi = query.request();
    const float *q = as_float_ptr(qi, dim(), "query", 1);
    PreparedQuery pq;
    pq.row.resize(info_.prepared_stride);
    float *o = pq.row.data();
    run([&] { return tq_prepare_queries(h_, q, 1, o, TQ_MEM_HOST, nullptr); });
    return pq;
  }

  // bindings.cpp:276-282
  PreparedSymCode prepare_symmetric_query(py::buffer code) {
    auto ci = code.request();
    const auto *c = static_cast<const uint8_t *>(as_bytes_ptr(ci, code_size(), "code"));
    PreparedSymCode pq;
    pq.row.resize(info_.sym_prepared_stride);
    float *o = pq.row.data();
    run([&] { return tq_prepare_symmetric(h_, c, 1, o, TQ_MEM_HOST, nullptr); });
    return pq;
  }

  // bindings.cpp:284-291
  float distance_symmetric_full_prepared(const PreparedSymCode &pq, py::buffer code_b) {
    check_sym(pq);
    auto ci = code_b.request();
    const auto *c = static_cast<const uint8_t *>(as_bytes_ptr(ci, code_size(), "code_b"));
    float d = 0.0f;
    const float *r = pq.row.data();
    run([&] { return tq_distance_symmetric_prepared(h_, r, 1, c, 1, &d, TQ_MEM_HOST, nullptr); });
    return d;
  }

  // bindings.cpp:293-310
  py::array_t<float> distance_1_to_n_symmetric_full(const PreparedSymCode &pq, py::buffer codes) {
    check_sym(pq);
    auto ci = codes.request();
    const size_t n = code_count(ci);
    py::array_t<float> out(static_cast<ssize_t>(n));
    float *o = out.mutable_data();
    const auto *c = static_cast<const uint8_t *>(ci.ptr);
    const float *r = pq.row.data();
    run([&] { return tq_distance_symmetric_prepared(h_, r, 1, c, n, o, TQ_MEM_HOST, nullptr); });
    return out;
  }

  // bindings.cpp:312-334: each a code prepared once, then scored against all b.
  py::array_t<float> distance_m_to_n_symmetric_full_prepared(py::buffer codes_a,
                                                             py::buffer codes_b) {
    return sym_matrix(codes_a, codes_b, TQ_SYM_FULL, TQ_MODE_FP32);
  }

 private:
  // bindings.cpp:159-167: itemsize 1, total bytes a multiple of code size.
  size_t code_count(const py::buffer_info &ci) const {
    if (ci.itemsize != 1) throw py::value_error("codes must be uint8");
    const ssize_t cs = static_cast<ssize_t>(code_size());
    ssize_t total = total_of(ci);
    if (total % cs != 0) throw py::value_error("codes size not a multiple of code_size_bytes");
    return static_cast<size_t>(total / cs);
  }

  // bindings.cpp:199-211 (and the full/light variants): both inputs itemsize
  // 1, totals multiples of the code size.
  py::array_t<float> sym_matrix(py::buffer codes_a, py::buffer codes_b, int variant, int md) {
    auto ai = codes_a.request();
    auto bi = codes_b.request();
    if (ai.itemsize != 1 || bi.itemsize != 1) throw py::value_error("codes must be uint8");
    const ssize_t cs = static_cast<ssize_t>(code_size());
    ssize_t ta = total_of(ai), tb = total_of(bi);
    if (ta % cs != 0 || tb % cs != 0)
      throw py::value_error("codes size not a multiple of code_size_bytes");
    const size_t m = ta / cs, n = tb / cs;
    py::array_t<float> out({static_cast<ssize_t>(m), static_cast<ssize_t>(n)});
    float *o = out.mutable_data();
    const auto *a = static_cast<const uint8_t *>(ai.ptr);
    const auto *b = static_cast<const uint8_t *>(bi.ptr);
    run([&] {
      return tq_distance_m_to_n_symmetric_ex(h_, a, m, b, n, o, variant, md, TQ_MEM_HOST,
                                             nullptr);
    });
    return out;
  }

  void check_sym(const PreparedSymCode &pq) const {
    if (pq.row.size() != info_.sym_prepared_stride)
      throw py::value_error("pq: prepared for a space of another dim");
  }

  template <class F>
  void run(F &&f) {
    int rc;
    {
      py::gil_scoped_release nogil;
      rc = f();
    }
    check(rc);
  }

  tq_space h_ = nullptr;
  tq_space_info info_{};
  int num_threads_;
};

}  // namespace

PYBIND11_MODULE(_turboquant, m) {
  m.doc() = "TurboQuant encode + quantized distances on AMD MI355X (gfx950)";
  py::class_<PreparedQuery>(m, "PreparedQuery");
  py::class_<PreparedSymCode>(m, "PreparedSymCode");
  py::class_<Space>(m, "TurboQuantSpace")
      .def(py::init<size_t, int, uint64_t, uint64_t, int>(), py::arg("dim"),
           py::arg("bits_per_coord") = 4, py::arg("rot_seed") = 42ULL,
           py::arg("qjl_seed") = 137ULL, py::arg("num_threads") = 0)
      .def("dim", &Space::dim)
      .def("padded_dim", &Space::padded_dim)
      .def("padded", &Space::padded)
      .def("num_threads", &Space::num_threads)
      .def("code_size_bytes", &Space::code_size)
      .def("bits_per_coord", &Space::bits)
      .def("encode", &Space::encode, py::arg("x"))
      .def("encode_into", &Space::encode_into, py::arg("x"), py::arg("out"))
      .def("encode_batch", &Space::encode_batch, py::arg("X"))
      .def("encode_batch_into", &Space::encode_batch_into, py::arg("X"), py::arg("out"))
      .def("distance", &Space::distance, py::arg("query"), py::arg("code"))
      .def("distance_symmetric", &Space::distance_symmetric, py::arg("code_a"),
           py::arg("code_b"))
      .def("distance_1_to_n", &Space::distance_1_to_n, py::arg("query"), py::arg("codes"))
      .def("distance_m_to_n", &Space::distance_m_to_n, py::arg("queries"), py::arg("codes"),
           py::arg("mode") = "fp32")
      .def("distance_m_to_n_symmetric", &Space::distance_m_to_n_symmetric, py::arg("codes_a"),
           py::arg("codes_b"), py::arg("mode") = "fp32")
      .def("distance_m_to_n_symmetric_full", &Space::distance_m_to_n_symmetric_full,
           py::arg("codes_a"), py::arg("codes_b"))
      .def("distance_m_to_n_symmetric_light", &Space::distance_m_to_n_symmetric_light,
           py::arg("codes_a"), py::arg("codes_b"))
      .def("rotated_coords_batch", &Space::rotated_coords_batch, py::arg("X"))
      .def("prepare_query", &Space::prepare_query, py::arg("query"))
      .def("prepare_symmetric_query", &Space::prepare_symmetric_query, py::arg("code"))
      .def("distance_symmetric_full_prepared", &Space::distance_symmetric_full_prepared,
           py::arg("pq"), py::arg("code_b"))
      .def("distance_1_to_n_symmetric_full", &Space::distance_1_to_n_symmetric_full,
           py::arg("pq"), py::arg("codes"))
      .def("distance_m_to_n_symmetric_full_prepared",
           &Space::distance_m_to_n_symmetric_full_prepared, py::arg("codes_a"), py::arg("codes_b"));
  m.def("device_count", &tq_device_count);
}
