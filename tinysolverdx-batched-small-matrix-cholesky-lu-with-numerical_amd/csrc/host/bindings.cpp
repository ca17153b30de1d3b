// pybind11 module `tinysolverdx._core`: NumPy arrays <-> the C++ host API.
//
// The module offers the entry points of the reference's src/python/bindings.cpp
// (names, argument names, defaults, exception types; ValueError is
// std::invalid_argument, RuntimeError std::runtime_error) with "hip" added to
// the backend names, has_hip(), and the fused posv / gesv drivers, whose
// counterparts the reference composes from two calls.  Each call works on
// fresh output arrays (inputs are never written) and releases the GIL while
// the device works.  The machine probes measure the GPU.  The *_device
// functions (device_arrays.cpp) take arrays already in GPU memory.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tinysolverdx/tinysolverdx.hpp"

namespace py = pybind11;

namespace {

// ------------------------------------------------------------ name tables
template <class E>
struct Named {
  const char* name;
  E value;
};
constexpr Named<tsdx::Precision> kPrecisionNames[] = {
    {"fp32", tsdx::Precision::FP32}, {"fp64", tsdx::Precision::FP64},
    {"fp16", tsdx::Precision::FP16}, {"bf16", tsdx::Precision::BF16}};
constexpr Named<tsdx::Backend> kBackendNames[] = {
    {"auto", tsdx::Backend::Auto},     {"hip", tsdx::Backend::Hip},   {"naive", tsdx::Backend::Naive},
    {"scalar", tsdx::Backend::Scalar}, {"simd", tsdx::Backend::Simd}, {"cuda", tsdx::Backend::Cuda}};

// The entry of `table` called `key`; an unknown key is a ValueError that lists
// the accepted names.
template <class E, size_t K>
E by_name(const Named<E> (&table)[K], const std::string& key, const char* what) {
  std::string names;
  for (const auto& e : table) {
    if (key == e.name) return e.value;
    names += names.empty() ? "" : ", ";
    names += e.name;
  }
  throw std::invalid_argument(std::string(what) + " must be one of " + names);
}
tsdx::Precision precision_of(const std::string& s) { return by_name(kPrecisionNames, s, "precision"); }
tsdx::Backend backend_of(const std::string& s) { return by_name(kBackendNames, s, "backend"); }

// ------------------------------------------------------------ array views
template <class S>
using Arr = py::array_t<S, py::array::c_style>;
using Ints = py::array_t<int, py::array::c_style>;
using Dbl = py::array_t<double, py::array::c_style>;

// A C-contiguous (batch, n, n) stack of matrices.
struct Dims {
  int n = 0;
  py::ssize_t batch = 0;
};
template <class S>
Dims square_stack(const Arr<S>& A, const char* what = "expected array of shape (batch, n, n)") {
  if (A.ndim() != 3 || A.shape(1) != A.shape(2)) throw std::invalid_argument(what);
  return Dims{static_cast<int>(A.shape(1)), A.shape(0)};
}
// (batch, n, k) right-hand sides matching d; returns k.
template <class S>
int rhs_of(const Arr<S>& B, const Dims& d) {
  if (B.ndim() != 3 || B.shape(0) != d.batch || B.shape(1) != d.n)
    throw std::invalid_argument("B must be (batch, n, nrhs)");
  return static_cast<int>(B.shape(2));
}
// A new array with x's shape and contents (outputs never alias inputs).
template <class S>
Arr<S> duplicate(const Arr<S>& x) {
  Arr<S> y(std::vector<py::ssize_t>(x.shape(), x.shape() + x.ndim()));
  if (x.size()) std::memcpy(y.mutable_data(), x.data(), sizeof(S) * static_cast<size_t>(x.size()));
  return y;
}
Ints int_array(std::vector<py::ssize_t> shape) { return Ints(std::move(shape)); }

// Runs f without the GIL.
template <class F>
void unlocked(F&& f) {
  py::gil_scoped_release release;
  f();
}

// ------------------------------------------------- per-dtype entry points
// Registers the solver entry points for storage type S under the suffix sfx
// (cholesky_f32, lu_f64, ...).
template <class S>
void add_solvers(py::module_& m, const std::string& sfx) {
  auto name = [&](const char* op) { return std::string(op) + "_" + sfx; };

  m.def(
      name("cholesky").c_str(),
      [](Arr<S> A, const std::string& prec, const std::string& be) {
        const Dims d = square_stack(A);
        Arr<S> L = duplicate(A);
        Ints info = int_array({d.batch});
        const auto p = precision_of(prec);
        const auto b = backend_of(be);
        unlocked([&] { tsdx::cholesky<S>(p, L.mutable_data(), d.n, d.batch, info.mutable_data(), b); });
        return py::make_tuple(L, info);
      },
      py::arg("A"), py::arg("precision") = "fp32", py::arg("backend") = "auto");

  m.def(
      name("lu").c_str(),
      [](Arr<S> A, const std::string& prec, const std::string& be) {
        const Dims d = square_stack(A);
        Arr<S> LU = duplicate(A);
        Ints piv = int_array({d.batch, d.n});
        Ints info = int_array({d.batch});
        const auto p = precision_of(prec);
        const auto b = backend_of(be);
        unlocked([&] {
          tsdx::lu<S>(p, LU.mutable_data(), d.n, d.batch, piv.mutable_data(), info.mutable_data(), b);
        });
        return py::make_tuple(LU, piv, info);
      },
      py::arg("A"), py::arg("precision") = "fp32", py::arg("backend") = "auto");

  m.def(
      name("cholesky_solve").c_str(),
      [](Arr<S> L, Arr<S> B, const std::string& prec) {
        const Dims d = square_stack(L);
        const int nrhs = rhs_of(B, d);
        Arr<S> X = duplicate(B);
        const auto p = precision_of(prec);
        unlocked([&] { tsdx::cholesky_solve<S>(p, L.data(), d.n, d.batch, X.mutable_data(), nrhs); });
        return X;
      },
      py::arg("L"), py::arg("B"), py::arg("precision") = "fp32");

  m.def(
      name("lu_solve").c_str(),
      [](Arr<S> LU, Ints piv, Arr<S> B, const std::string& prec) {
        const Dims d = square_stack(LU);
        const int nrhs = rhs_of(B, d);
        if (piv.ndim() != 2 || piv.shape(0) != d.batch || piv.shape(1) != d.n)
          throw std::invalid_argument("piv must be (batch, n)");
        Arr<S> X = duplicate(B);
        const auto p = precision_of(prec);
        unlocked([&] { tsdx::lu_solve<S>(p, LU.data(), piv.data(), d.n, d.batch, X.mutable_data(), nrhs); });
        return X;
      },
      py::arg("LU"), py::arg("piv"), py::arg("B"), py::arg("precision") = "fp32");

  // fused drivers (one device launch): (L, X, info) and (LU, piv, X, info)
  m.def(
      name("posv").c_str(),
      [](Arr<S> A, Arr<S> B, const std::string& prec, const std::string& be) {
        const Dims d = square_stack(A);
        const int nrhs = rhs_of(B, d);
        Arr<S> L = duplicate(A), X = duplicate(B);
        Ints info = int_array({d.batch});
        const auto p = precision_of(prec);
        const auto b = backend_of(be);
        unlocked([&] {
          tsdx::posv<S>(p, L.mutable_data(), d.n, d.batch, X.mutable_data(), nrhs, info.mutable_data(), b);
        });
        return py::make_tuple(L, X, info);
      },
      py::arg("A"), py::arg("B"), py::arg("precision") = "fp32", py::arg("backend") = "auto");

  m.def(
      name("gesv").c_str(),
      [](Arr<S> A, Arr<S> B, const std::string& prec, const std::string& be) {
        const Dims d = square_stack(A);
        const int nrhs = rhs_of(B, d);
        Arr<S> LU = duplicate(A), X = duplicate(B);
        Ints piv = int_array({d.batch, d.n});
        Ints info = int_array({d.batch});
        const auto p = precision_of(prec);
        const auto b = backend_of(be);
        unlocked([&] {
          tsdx::gesv<S>(p, LU.mutable_data(), d.n, d.batch, X.mutable_data(), nrhs, piv.mutable_data(),
                        info.mutable_data(), b);
        });
        return py::make_tuple(LU, piv, X, info);
      },
      py::arg("A"), py::arg("B"), py::arg("precision") = "fp32", py::arg("backend") = "auto");

  m.def(
      name("round_to").c_str(),
      [](Arr<S> x, const std::string& prec) {
        Arr<S> y = duplicate(x);
        tsdx::round_to<S>(precision_of(prec), y.mutable_data(), y.size());
        return y;
      },
      py::arg("x"), py::arg("precision"));
}

// ------------------------------------------------------------- metrics
// Every metric maps float64 inputs to one value per matrix.
using PerMatrix = void (*)(const double*, const double*, const double*, int, std::int64_t, double*);

// (batch, n) vectors matching the matrices.
void need_vectors(const Dims& d, const Dbl& v, const char* what) {
  if (v.ndim() != 2 || v.shape(0) != d.batch || v.shape(1) != d.n)
    throw std::invalid_argument(std::string(what) + " must be (batch, n)");
}

void add_metrics(py::module_& m) {
  m.def("chol_backward_error", [](Dbl A, Dbl L) {
    if (A.ndim() != 3 || A.shape(1) != A.shape(2) || L.ndim() != 3 || L.size() != A.size())
      throw std::invalid_argument("shape mismatch");
    const Dims d = square_stack(A);
    Dbl out(std::vector<py::ssize_t>{d.batch});
    unlocked([&] { tsdx::chol_backward_error(A.data(), L.data(), d.n, d.batch, out.mutable_data()); });
    return out;
  });
  m.def("lu_backward_error", [](Dbl A, Dbl LU, Ints piv) {
    if (A.ndim() != 3 || A.shape(1) != A.shape(2) || LU.size() != A.size() || piv.size() != A.shape(0) * A.shape(1))
      throw std::invalid_argument("shape mismatch");
    const Dims d = square_stack(A);
    Dbl out(std::vector<py::ssize_t>{d.batch});
    unlocked([&] { tsdx::lu_backward_error(A.data(), LU.data(), piv.data(), d.n, d.batch, out.mutable_data()); });
    return out;
  });
  // (A, x, b) -> one value per matrix
  const std::pair<const char*, PerMatrix> vector_metrics[] = {
      {"solve_backward_error", &tsdx::solve_backward_error},
      {"solve_forward_error", &tsdx::solve_forward_error},
  };
  for (const auto& vm : vector_metrics) {
    const PerMatrix f = vm.second;
    m.def(vm.first, [f](Dbl A, Dbl x, Dbl b) {
      const Dims d = square_stack(A, "A must be (batch, n, n)");
      need_vectors(d, x, "x");
      need_vectors(d, b, "b");
      Dbl out(std::vector<py::ssize_t>{d.batch});
      unlocked([&] { f(A.data(), x.data(), b.data(), d.n, d.batch, out.mutable_data()); });
      return out;
    });
  }
  m.def("reference_solve", [](Dbl A, Dbl b) {
    const Dims d = square_stack(A, "A must be (batch, n, n)");
    need_vectors(d, b, "b");
    Dbl x(std::vector<py::ssize_t>{d.batch, d.n});
    unlocked([&] { tsdx::reference_solve(A.data(), b.data(), d.n, d.batch, x.mutable_data()); });
    return x;
  });
}

}  // namespace

namespace tsdx_py {
void add_device_arrays(py::module_& m);  // device_arrays.cpp
}

PYBIND11_MODULE(_core, m) {
  m.doc() = "tinysolverdx-hip native core: batched small-matrix Cholesky/LU on MI355X (gfx950)";
  add_solvers<float>(m, "f32");
  add_solvers<double>(m, "f64");
  add_metrics(m);
  tsdx_py::add_device_arrays(m);

  m.def("has_hip", &tsdx::has_hip);
  m.def("has_cuda", &tsdx::has_cuda);
  m.def("has_openmp", [] { return false; });  // host threads are std::thread
  m.def("num_threads", &tsdx::num_threads);
  m.def("set_num_threads", &tsdx::set_num_threads);
  m.def("hip_version", [] { return tsdx_hip_version(); });
  m.def("unit_roundoff", [](const std::string& p) { return tsdx::unit_roundoff(precision_of(p)); });

  // machine probes of the GPU (the reference measures its CPU here)
  m.def("measure_peak_gflops", [](const std::string& p) {
    const auto prec = precision_of(p);
    double r = 0.0;
    unlocked([&] { r = tsdx::measure_peak_gflops(prec); });
    return r;
  });
  m.def("measure_bandwidth_gbs", [] {
    double r = 0.0;
    unlocked([&] { r = tsdx::measure_bandwidth_gbs(); });
    return r;
  });
  m.def(
      "measure_mfma_tflops",
      [](const std::string& p) {
        const auto prec = precision_of(p);
        double r = 0.0;
        unlocked([&] { r = tsdx::measure_mfma_tflops(prec); });
        return r;
      },
      py::arg("precision"));
}
