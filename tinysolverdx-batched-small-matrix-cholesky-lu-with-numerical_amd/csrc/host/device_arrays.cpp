// Device arrays in the Python module: the solvers on matrices that already
// live in GPU memory, with no trip over PCIe (the reference's API takes host
// arrays only and stages them, README.md:99-101).
//
// Inputs: any object that exports __cuda_array_interface__ (version 2 or 3) or
// DLPack (__dlpack__ / __dlpack_device__ on a ROCm or CUDA device): torch
// tensors, CuPy / Numba arrays, and this module's DeviceArray.  They must be
// C-contiguous float32 / float64 (int32 pivots) on the current device.
// Outputs: DeviceArray, memory owned by the library that exports both
// protocols, so torch.from_dlpack / cupy.asarray / numba take it without a
// copy; DeviceArray.numpy() copies to the host.
//
// Every call synchronises the device before it reads its inputs (the
// producer's work on any stream is complete) and after its last launch (the
// results are ready when it returns), like the host-array calls.  Inputs are
// never written.  Precision: the array's own (float32: fp32, float64: fp64) or
// the emulated fp16 / bf16 on either storage.
#include <Python.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dlpack_abi.hpp"
#include "tinysolverdx/tinysolverdx.hpp"

namespace py = pybind11;

namespace tsdx_py {
namespace {

// ---------------------------------------------------------- element types
enum class Dt { F32, F64, I32 };

size_t item_size(Dt t) { return t == Dt::F64 ? 8 : 4; }
const char* type_str(Dt t) { return t == Dt::F32 ? "<f4" : t == Dt::F64 ? "<f8" : "<i4"; }
const char* numpy_name(Dt t) { return t == Dt::F32 ? "float32" : t == Dt::F64 ? "float64" : "int32"; }

void check(int rc) {
  if (rc == TSDX_HIP_OK) return;
  const std::string msg = tsdx_hip_last_error();
  if (rc == TSDX_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error(msg.empty() ? std::string("HIP backend failure") : msg);
}

int64_t count_of(const std::vector<int64_t>& shape) {
  int64_t c = 1;
  for (int64_t s : shape) c *= s;
  return c;
}

// ------------------------------------------------------------ DeviceArray
struct DeviceArray {
  void* ptr = nullptr;
  Dt dt = Dt::F32;
  std::vector<int64_t> shape;
  int device = 0;

  DeviceArray(Dt t, std::vector<int64_t> s) : dt(t), shape(std::move(s)) {
    check(tsdx_hip_get_device(&device));
    check(tsdx_hip_malloc(&ptr, bytes()));
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray() {
    if (ptr) tsdx_hip_free(ptr);
  }
  size_t bytes() const { return item_size(dt) * static_cast<size_t>(count_of(shape)); }
};
using DevPtr = std::shared_ptr<DeviceArray>;

py::tuple shape_tuple(const std::vector<int64_t>& shape) {
  py::tuple t(shape.size());
  for (size_t i = 0; i < shape.size(); ++i) t[i] = py::int_(shape[i]);
  return t;
}

// DLPack export: the managed tensor owns a reference to the array (a heap
// copy of the shared pointer) and the shape; its deleter needs no Python.
struct Exported {
  DLManagedTensor mt;
  std::vector<int64_t> shape;
  DevPtr owner;
};
void exported_deleter(DLManagedTensor* self) { delete static_cast<Exported*>(self->manager_ctx); }
void capsule_destructor(PyObject* cap) {
  // a capsule never consumed still owns its tensor
  if (PyCapsule_IsValid(cap, "dltensor")) {
    auto* mt = static_cast<DLManagedTensor*>(PyCapsule_GetPointer(cap, "dltensor"));
    if (mt && mt->deleter) mt->deleter(mt);
  }
}

py::object to_dlpack(const DevPtr& a) {
  auto* e = new Exported();
  e->shape = a->shape;
  e->owner = a;
  e->mt.manager_ctx = e;
  e->mt.deleter = &exported_deleter;
  DLTensor& t = e->mt.dl_tensor;
  t.data = a->ptr;
  t.device = DLDevice{kDLROCM, a->device};
  t.ndim = static_cast<int32_t>(a->shape.size());
  t.dtype = DLDataType{a->dt == Dt::I32 ? kDLInt : kDLFloat, static_cast<uint8_t>(8 * item_size(a->dt)), 1};
  t.shape = e->shape.data();
  t.strides = nullptr;
  t.byte_offset = 0;
  PyObject* cap = PyCapsule_New(&e->mt, "dltensor", &capsule_destructor);
  if (!cap) {
    delete e;
    throw py::error_already_set();
  }
  return py::reinterpret_steal<py::object>(cap);
}

// ------------------------------------------------------------ input views
// A device input: pointer, element type and shape; a DLPack import keeps the
// producer's managed tensor until the view is destroyed.
struct DevIn {
  void* ptr = nullptr;
  Dt dt = Dt::F32;
  std::vector<int64_t> shape;
  DLManagedTensor* mt = nullptr;
  DevIn() = default;
  DevIn(const DevIn&) = delete;
  DevIn(DevIn&& o) noexcept : ptr(o.ptr), dt(o.dt), shape(std::move(o.shape)), mt(o.mt) { o.mt = nullptr; }
  ~DevIn() {
    if (mt && mt->deleter) mt->deleter(mt);
  }
  size_t bytes() const { return item_size(dt) * static_cast<size_t>(count_of(shape)); }
};

Dt dt_of_typestr(const std::string& s, const char* what) {
  if (s == "<f4") return Dt::F32;
  if (s == "<f8") return Dt::F64;
  if (s == "<i4") return Dt::I32;
  throw py::type_error(std::string(what) + ": device arrays must be float32 or float64 (int32 pivots), got " + s);
}

Dt dt_of_dlpack(DLDataType t, const char* what) {
  if (t.lanes == 1 && t.code == kDLFloat && t.bits == 32) return Dt::F32;
  if (t.lanes == 1 && t.code == kDLFloat && t.bits == 64) return Dt::F64;
  if (t.lanes == 1 && t.code == kDLInt && t.bits == 32) return Dt::I32;
  throw py::type_error(std::string(what) + ": device arrays must be float32 or float64 (int32 pivots)");
}

// Contiguous row-major strides (in bytes when item > 1 is given in bytes).
bool row_major(const std::vector<int64_t>& shape, const std::vector<int64_t>& strides, int64_t item) {
  int64_t expect = item;
  for (size_t k = shape.size(); k-- > 0;) {
    if (shape[k] != 1 && strides[k] != expect) return false;
    expect *= shape[k];
  }
  return true;
}

// The memory must be on the current device.
void check_on_device(const void* ptr, size_t bytes, const char* what) {
  if (bytes == 0) return;
  int dev = -1, cur = -1;
  const int rc = tsdx_hip_pointer_device(ptr, &dev);
  if (rc == TSDX_HIP_ERUNTIME) check(rc);  // no device at all
  if (rc != TSDX_HIP_OK) throw std::invalid_argument(std::string(what) + ": " + tsdx_hip_last_error());
  check(tsdx_hip_get_device(&cur));
  if (dev != cur)
    throw std::invalid_argument(std::string(what) + ": array is on device " + std::to_string(dev) +
                                ", the current device is " + std::to_string(cur));
}

bool has_cai(py::handle o) { return PyObject_HasAttrString(o.ptr(), "__cuda_array_interface__") == 1; }

bool dlpack_on_gpu(py::handle o) {
  if (!py::hasattr(o, "__dlpack__") || !py::hasattr(o, "__dlpack_device__")) return false;
  py::tuple d = o.attr("__dlpack_device__")();
  const int t = d[0].cast<int>();
  return t == kDLROCM || t == kDLCUDA;
}

DevIn view_of(py::handle o, const char* what) {
  DevIn v;
  if (has_cai(o)) {
    py::dict cai = o.attr("__cuda_array_interface__");
    if (cai.contains("mask") && !cai["mask"].is_none())
      throw std::invalid_argument(std::string(what) + ": masked arrays are not supported");
    v.dt = dt_of_typestr(cai["typestr"].cast<std::string>(), what);
    for (auto s : cai["shape"].cast<py::tuple>()) v.shape.push_back(s.cast<int64_t>());
    py::tuple data = cai["data"].cast<py::tuple>();
    v.ptr = reinterpret_cast<void*>(data[0].cast<uintptr_t>());
    if (cai.contains("strides") && !cai["strides"].is_none()) {
      std::vector<int64_t> st;
      for (auto s : cai["strides"].cast<py::tuple>()) st.push_back(s.cast<int64_t>());
      if (!row_major(v.shape, st, static_cast<int64_t>(item_size(v.dt))))
        throw std::invalid_argument(std::string(what) + ": device arrays must be C-contiguous");
    }
  } else if (dlpack_on_gpu(o)) {
    py::object cap = o.attr("__dlpack__")();
    auto* mt = static_cast<DLManagedTensor*>(PyCapsule_GetPointer(cap.ptr(), "dltensor"));
    if (!mt) throw py::error_already_set();
    PyCapsule_SetName(cap.ptr(), "used_dltensor");  // consumed: the view calls the deleter
    v.mt = mt;
    const DLTensor& t = mt->dl_tensor;
    v.dt = dt_of_dlpack(t.dtype, what);
    v.shape.assign(t.shape, t.shape + t.ndim);
    if (t.strides) {
      std::vector<int64_t> st(t.strides, t.strides + t.ndim);
      if (!row_major(v.shape, st, 1))
        throw std::invalid_argument(std::string(what) + ": device arrays must be C-contiguous");
    }
    v.ptr = static_cast<char*>(t.data) + t.byte_offset;
  } else {
    throw py::type_error(std::string(what) + ": not a device array (no __cuda_array_interface__ or GPU DLPack)");
  }
  check_on_device(v.ptr, v.bytes(), what);
  return v;
}

// ------------------------------------------------------- shapes and modes
struct Mats {
  int n = 0;
  int64_t batch = 0;
  bool single = false;
};
Mats mats_of(const DevIn& a, const char* what) {
  if (a.dt == Dt::I32) throw py::type_error(std::string(what) + " must be float32 or float64");
  if (a.shape.size() == 2 && a.shape[0] == a.shape[1]) return Mats{static_cast<int>(a.shape[0]), 1, true};
  if (a.shape.size() == 3 && a.shape[1] == a.shape[2])
    return Mats{static_cast<int>(a.shape[1]), a.shape[0], false};
  throw std::invalid_argument(std::string(what) + " must be (n, n) or (batch, n, n)");
}

// Right-hand sides for m: one matrix takes b (n,) or B (n, k), a stack
// (batch, n) or (batch, n, k).  Returns k.
int rhs_of(const DevIn& b, const Mats& m, const DevIn& a) {
  if (b.dt != a.dt) throw py::type_error("B must have the dtype of the matrices");
  const size_t vec = m.single ? 1 : 2;
  const auto& s = b.shape;
  const bool lead_ok = m.single || (!s.empty() && s[0] == m.batch);
  if (lead_ok && s.size() == vec && s[vec - 1] == m.n) return 1;
  if (lead_ok && s.size() == vec + 1 && s[vec - 1] == m.n) return static_cast<int>(s[vec]);
  throw std::invalid_argument(m.single ? "B must be (n,) or (n, k)" : "B must be (batch, n) or (batch, n, k)");
}

// Resolves the precision: "" is the array's own; the emulated formats give a
// TSDX_HIP_EMU_* code, the native ones 0.
int emu_code(const std::string& prec, Dt dt) {
  if (prec.empty() || (prec == "fp32" && dt == Dt::F32) || (prec == "fp64" && dt == Dt::F64)) return 0;
  if (prec == "fp16") return TSDX_HIP_EMU_FP16;
  if (prec == "bf16") return TSDX_HIP_EMU_BF16;
  if (prec == "fp32" || prec == "fp64")
    throw std::invalid_argument("on device arrays the precision is the array's own (float32: fp32, float64: "
                                "fp64) or the emulated fp16 / bf16");
  throw std::invalid_argument("precision must be one of fp16, bf16, fp32, fp64");
}

void require_hip_backend(const std::string& be) {
  if (be == "auto" || be == "hip") return;
  if (be == "cuda") throw std::runtime_error("CUDA backend unavailable for this build/precision");
  if (be == "naive" || be == "scalar" || be == "simd")
    throw std::runtime_error("backend '" + be +
                             "' is a CPU backend of the reference library and is not part of tinysolverdx-hip");
  throw std::invalid_argument("backend must be one of auto, hip, naive, scalar, simd, cuda");
}

std::vector<int64_t> batch_shape(const Mats& m) { return m.single ? std::vector<int64_t>{} : std::vector<int64_t>{m.batch}; }
std::vector<int64_t> piv_shape(const Mats& m) {
  return m.single ? std::vector<int64_t>{m.n} : std::vector<int64_t>{m.batch, m.n};
}

// Device work without the GIL, framed by device synchronisations.
template <class F>
void run(F&& f) {
  py::gil_scoped_release release;
  check(tsdx_hip_device_synchronize());
  f();
  check(tsdx_hip_device_synchronize());
}
void copy(void* dst, const void* src, size_t bytes) { check(tsdx_hip_memcpy(dst, src, bytes)); }

// ------------------------------------------------------------- operations
py::tuple cholesky_device(py::handle A, const std::string& prec, const std::string& be) {
  require_hip_backend(be);
  DevIn a = view_of(A, "A");
  const Mats m = mats_of(a, "A");
  const int emu = emu_code(prec, a.dt);
  auto L = std::make_shared<DeviceArray>(a.dt, a.shape);
  auto info = std::make_shared<DeviceArray>(Dt::I32, batch_shape(m));
  run([&] {
    copy(L->ptr, a.ptr, a.bytes());
    int* ip = static_cast<int*>(info->ptr);
    if (a.dt == Dt::F32)
      check(emu ? tsdx_hip_potrf_emu_f32(static_cast<float*>(L->ptr), m.n, m.batch, ip, emu, nullptr)
                : tsdx_hip_potrf_f32(static_cast<float*>(L->ptr), m.n, m.batch, ip, nullptr));
    else
      check(emu ? tsdx_hip_potrf_emu_f64(static_cast<double*>(L->ptr), m.n, m.batch, ip, emu, nullptr)
                : tsdx_hip_potrf_f64(static_cast<double*>(L->ptr), m.n, m.batch, ip, nullptr));
  });
  return py::make_tuple(L, info);
}

py::tuple lu_device(py::handle A, const std::string& prec, const std::string& be) {
  require_hip_backend(be);
  DevIn a = view_of(A, "A");
  const Mats m = mats_of(a, "A");
  const int emu = emu_code(prec, a.dt);
  auto LU = std::make_shared<DeviceArray>(a.dt, a.shape);
  auto piv = std::make_shared<DeviceArray>(Dt::I32, piv_shape(m));
  auto info = std::make_shared<DeviceArray>(Dt::I32, batch_shape(m));
  run([&] {
    copy(LU->ptr, a.ptr, a.bytes());
    int* pp = static_cast<int*>(piv->ptr);
    int* ip = static_cast<int*>(info->ptr);
    if (a.dt == Dt::F32)
      check(emu ? tsdx_hip_getrf_emu_f32(static_cast<float*>(LU->ptr), m.n, m.batch, pp, ip, emu, nullptr)
                : tsdx_hip_getrf_f32(static_cast<float*>(LU->ptr), m.n, m.batch, pp, ip, nullptr));
    else
      check(emu ? tsdx_hip_getrf_emu_f64(static_cast<double*>(LU->ptr), m.n, m.batch, pp, ip, emu, nullptr)
                : tsdx_hip_getrf_f64(static_cast<double*>(LU->ptr), m.n, m.batch, pp, ip, nullptr));
  });
  return py::make_tuple(LU, piv, info);
}

DevPtr cholesky_solve_device(py::handle Lh, py::handle Bh, const std::string& prec) {
  DevIn l = view_of(Lh, "L");
  const Mats m = mats_of(l, "L");
  DevIn b = view_of(Bh, "B");
  const int k = rhs_of(b, m, l);
  const int emu = emu_code(prec, l.dt);
  auto X = std::make_shared<DeviceArray>(b.dt, b.shape);
  run([&] {
    copy(X->ptr, b.ptr, b.bytes());
    if (l.dt == Dt::F32)
      check(emu ? tsdx_hip_potrs_emu_f32(static_cast<const float*>(l.ptr), m.n, m.batch, static_cast<float*>(X->ptr),
                                         k, emu, nullptr)
                : tsdx_hip_potrs_f32(static_cast<const float*>(l.ptr), m.n, m.batch, static_cast<float*>(X->ptr), k,
                                     nullptr));
    else
      check(emu ? tsdx_hip_potrs_emu_f64(static_cast<const double*>(l.ptr), m.n, m.batch,
                                         static_cast<double*>(X->ptr), k, emu, nullptr)
                : tsdx_hip_potrs_f64(static_cast<const double*>(l.ptr), m.n, m.batch, static_cast<double*>(X->ptr),
                                     k, nullptr));
  });
  return X;
}

DevPtr lu_solve_device(py::handle LUh, py::handle Ph, py::handle Bh, const std::string& prec) {
  DevIn lu = view_of(LUh, "LU");
  const Mats m = mats_of(lu, "LU");
  DevIn p = view_of(Ph, "piv");
  if (p.dt != Dt::I32 || p.shape != piv_shape(m))
    throw std::invalid_argument(m.single ? "piv must be int32 (n,)" : "piv must be int32 (batch, n)");
  DevIn b = view_of(Bh, "B");
  const int k = rhs_of(b, m, lu);
  const int emu = emu_code(prec, lu.dt);
  auto X = std::make_shared<DeviceArray>(b.dt, b.shape);
  run([&] {
    copy(X->ptr, b.ptr, b.bytes());
    const int* pp = static_cast<const int*>(p.ptr);
    if (lu.dt == Dt::F32)
      check(emu ? tsdx_hip_getrs_emu_f32(static_cast<const float*>(lu.ptr), pp, m.n, m.batch,
                                         static_cast<float*>(X->ptr), k, emu, nullptr)
                : tsdx_hip_getrs_f32(static_cast<const float*>(lu.ptr), pp, m.n, m.batch, static_cast<float*>(X->ptr),
                                     k, nullptr));
    else
      check(emu ? tsdx_hip_getrs_emu_f64(static_cast<const double*>(lu.ptr), pp, m.n, m.batch,
                                         static_cast<double*>(X->ptr), k, emu, nullptr)
                : tsdx_hip_getrs_f64(static_cast<const double*>(lu.ptr), pp, m.n, m.batch,
                                     static_cast<double*>(X->ptr), k, nullptr));
  });
  return X;
}

// posv / gesv: (X, info).  The native precisions run the solve-only fused
// kernels straight from A and B (no copy, no stored factor); the emulated
// ones factor a device copy of A, then solve.
py::tuple posv_device(py::handle Ah, py::handle Bh, const std::string& prec, const std::string& be) {
  require_hip_backend(be);
  DevIn a = view_of(Ah, "A");
  const Mats m = mats_of(a, "A");
  DevIn b = view_of(Bh, "B");
  const int k = rhs_of(b, m, a);
  const int emu = emu_code(prec, a.dt);
  auto X = std::make_shared<DeviceArray>(b.dt, b.shape);
  auto info = std::make_shared<DeviceArray>(Dt::I32, batch_shape(m));
  run([&] {
    int* ip = static_cast<int*>(info->ptr);
    if (!emu) {
      if (a.dt == Dt::F32)
        check(tsdx_hip_posv_solve_f32(static_cast<const float*>(a.ptr), m.n, m.batch,
                                      static_cast<const float*>(b.ptr), static_cast<float*>(X->ptr), k, ip, nullptr));
      else
        check(tsdx_hip_posv_solve_f64(static_cast<const double*>(a.ptr), m.n, m.batch,
                                      static_cast<const double*>(b.ptr), static_cast<double*>(X->ptr), k, ip,
                                      nullptr));
      return;
    }
    DeviceArray L(a.dt, a.shape);
    copy(L.ptr, a.ptr, a.bytes());
    copy(X->ptr, b.ptr, b.bytes());
    if (a.dt == Dt::F32) {
      check(tsdx_hip_potrf_emu_f32(static_cast<float*>(L.ptr), m.n, m.batch, ip, emu, nullptr));
      check(tsdx_hip_potrs_emu_f32(static_cast<const float*>(L.ptr), m.n, m.batch, static_cast<float*>(X->ptr), k,
                                   emu, nullptr));
    } else {
      check(tsdx_hip_potrf_emu_f64(static_cast<double*>(L.ptr), m.n, m.batch, ip, emu, nullptr));
      check(tsdx_hip_potrs_emu_f64(static_cast<const double*>(L.ptr), m.n, m.batch, static_cast<double*>(X->ptr), k,
                                   emu, nullptr));
    }
    check(tsdx_hip_device_synchronize());  // L is freed at the end of this scope
  });
  return py::make_tuple(X, info);
}

py::tuple gesv_device(py::handle Ah, py::handle Bh, const std::string& prec, const std::string& be) {
  require_hip_backend(be);
  DevIn a = view_of(Ah, "A");
  const Mats m = mats_of(a, "A");
  DevIn b = view_of(Bh, "B");
  const int k = rhs_of(b, m, a);
  const int emu = emu_code(prec, a.dt);
  auto X = std::make_shared<DeviceArray>(b.dt, b.shape);
  auto info = std::make_shared<DeviceArray>(Dt::I32, batch_shape(m));
  run([&] {
    int* ip = static_cast<int*>(info->ptr);
    if (!emu) {
      if (a.dt == Dt::F32)
        check(tsdx_hip_gesv_solve_f32(static_cast<const float*>(a.ptr), m.n, m.batch,
                                      static_cast<const float*>(b.ptr), static_cast<float*>(X->ptr), k, nullptr, ip,
                                      nullptr));
      else
        check(tsdx_hip_gesv_solve_f64(static_cast<const double*>(a.ptr), m.n, m.batch,
                                      static_cast<const double*>(b.ptr), static_cast<double*>(X->ptr), k, nullptr,
                                      ip, nullptr));
      return;
    }
    DeviceArray LU(a.dt, a.shape), piv(Dt::I32, piv_shape(m));
    copy(LU.ptr, a.ptr, a.bytes());
    copy(X->ptr, b.ptr, b.bytes());
    int* pp = static_cast<int*>(piv.ptr);
    if (a.dt == Dt::F32) {
      check(tsdx_hip_getrf_emu_f32(static_cast<float*>(LU.ptr), m.n, m.batch, pp, ip, emu, nullptr));
      check(tsdx_hip_getrs_emu_f32(static_cast<const float*>(LU.ptr), pp, m.n, m.batch, static_cast<float*>(X->ptr),
                                   k, emu, nullptr));
    } else {
      check(tsdx_hip_getrf_emu_f64(static_cast<double*>(LU.ptr), m.n, m.batch, pp, ip, emu, nullptr));
      check(tsdx_hip_getrs_emu_f64(static_cast<const double*>(LU.ptr), pp, m.n, m.batch,
                                   static_cast<double*>(X->ptr), k, emu, nullptr));
    }
    check(tsdx_hip_device_synchronize());  // LU and piv are freed at the end of this scope
  });
  return py::make_tuple(X, info);
}

}  // namespace

void add_device_arrays(py::module_& m) {
  py::class_<DeviceArray, DevPtr>(m, "DeviceArray",
                                  "GPU memory owned by the library (a solver result); exports "
                                  "__cuda_array_interface__ and DLPack, numpy() copies to the host")
      .def_property_readonly("shape", [](const DeviceArray& a) { return shape_tuple(a.shape); })
      .def_property_readonly("dtype", [](const DeviceArray& a) { return py::dtype(numpy_name(a.dt)); })
      .def_property_readonly("ndim", [](const DeviceArray& a) { return a.shape.size(); })
      .def_property_readonly("size", [](const DeviceArray& a) { return count_of(a.shape); })
      .def_property_readonly("device", [](const DeviceArray& a) { return a.device; })
      .def_property_readonly("__cuda_array_interface__",
                             [](const DeviceArray& a) {
                               py::dict d;
                               d["shape"] = shape_tuple(a.shape);
                               d["typestr"] = type_str(a.dt);
                               d["data"] = py::make_tuple(reinterpret_cast<uintptr_t>(a.ptr), false);
                               d["strides"] = py::none();
                               d["version"] = 3;
                               d["stream"] = py::none();  // ready: every call synchronises before returning
                               return d;
                             })
      .def("__dlpack_device__", [](const DeviceArray& a) { return py::make_tuple(kDLROCM, a.device); })
      .def(
          "__dlpack__", [](const DevPtr& a, py::args, py::kwargs) { return to_dlpack(a); },
          "DLPack capsule of the array (no copy; the data are ready on every stream)")
      .def("numpy",
           [](const DeviceArray& a) {
             std::vector<py::ssize_t> shape(a.shape.begin(), a.shape.end());
             py::array out(py::dtype(numpy_name(a.dt)), shape);
             {
               py::gil_scoped_release release;
               check(tsdx_hip_memcpy(out.mutable_data(), a.ptr, a.bytes()));
             }
             return out;
           })
      .def("__len__",
           [](const DeviceArray& a) {
             if (a.shape.empty()) throw py::type_error("len() of a 0-d array");
             return a.shape[0];
           })
      .def("__repr__", [](const DeviceArray& a) {
        return "DeviceArray(shape=" + py::repr(shape_tuple(a.shape)).cast<std::string>() + ", dtype=" +
               numpy_name(a.dt) + ", device=" + std::to_string(a.device) + ")";
      });

  m.def("is_device_array", [](py::handle o) { return has_cai(o) || dlpack_on_gpu(o); });
  m.def("cholesky_device", &cholesky_device, py::arg("A"), py::arg("precision") = "", py::arg("backend") = "auto");
  m.def("lu_device", &lu_device, py::arg("A"), py::arg("precision") = "", py::arg("backend") = "auto");
  m.def("cholesky_solve_device", &cholesky_solve_device, py::arg("L"), py::arg("B"), py::arg("precision") = "");
  m.def("lu_solve_device", &lu_solve_device, py::arg("LU"), py::arg("piv"), py::arg("B"), py::arg("precision") = "");
  m.def("posv_device", &posv_device, py::arg("A"), py::arg("B"), py::arg("precision") = "",
        py::arg("backend") = "auto");
  m.def("gesv_device", &gesv_device, py::arg("A"), py::arg("B"), py::arg("precision") = "",
        py::arg("backend") = "auto");
}

}  // namespace tsdx_py
