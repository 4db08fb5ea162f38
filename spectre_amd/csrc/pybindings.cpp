// pybind11 module spectre_amd._Pybindings: the Python face of the device path,
// mirroring the reference's bindings
//   spectre.Spectral                           (Mesh, Basis, Quadrature,
//                                               collocation_points;
//                                               NumericalAlgorithms/Spectral/Python/)
//   spectre.NumericalAlgorithms.LinearOperators (partial_derivative,
//                                               logical_partial_derivative;
//                                               LinearOperators/Python/PartialDerivatives.cpp:20-40)
// plus the device right-hand sides, their adjoints and the RK3 substeps on a
// domain handle, which the torch.autograd Functions of the package call.
//
// Tensors are torch tensors on the HIP device (float64, contiguous), reached
// through their Python attributes (data_ptr, shape, ...), so the module does
// not link libtorch; outputs are allocated with torch.empty on the input's
// device and work is enqueued on torch's current stream.  Every call goes
// through the C ABI of libspectre_amd.so; a non-zero status raises
// RuntimeError with the library's message.  Host code only.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spectre_amd.h"

namespace py = pybind11;

namespace {

enum class Basis { Legendre, Chebyshev, FiniteDifference, SphericalHarmonic };
enum class Quadrature { GaussLobatto, Gauss, CellCentered, FaceCentered };

// Mesh<Dim> (NumericalAlgorithms/Spectral/Mesh.hpp:53-243): extents per
// dimension with one basis and quadrature.  Only Legendre / Gauss-Lobatto
// meshes have device kernels.
template <int Dim>
struct MeshD {
  std::array<int, Dim> ext{};
  Basis basis = Basis::Legendre;
  Quadrature quadrature = Quadrature::GaussLobatto;
  int64_t number_of_grid_points() const {
    int64_t n = 1;
    for (int d = 0; d < Dim; ++d) n *= ext[d];
    return n;
  }
};

void check(int status) {
  if (status != SXD_OK) throw std::runtime_error(std::string("spectre_amd: ") + sxd_last_error());
}

py::object torch_module() { return py::module_::import("torch"); }

// Device pointer of a float64 contiguous device tensor, after checking it.
void* device_ptr(const py::object& t, const char* name) {
  py::object torch = torch_module();
  if (!py::isinstance(t, torch.attr("Tensor")))
    throw py::type_error(std::string(name) + " must be a torch.Tensor");
  if (!t.attr("is_cuda").cast<bool>())
    throw py::value_error(std::string(name) + " must be a device tensor");
  if (!t.attr("dtype").equal(torch.attr("float64")))
    throw py::value_error(std::string(name) + " must be float64");
  if (!t.attr("is_contiguous")().cast<bool>())
    throw py::value_error(std::string(name) + " must be contiguous");
  return reinterpret_cast<void*>(t.attr("data_ptr")().cast<uintptr_t>());
}

std::vector<int64_t> shape_of(const py::object& t) {
  return t.attr("shape").cast<std::vector<int64_t>>();
}

void* current_stream(const py::object& t) {
  py::object s = torch_module().attr("cuda").attr("current_stream")(t.attr("device"));
  return reinterpret_cast<void*>(s.attr("cuda_stream").cast<uintptr_t>());
}

py::object empty_like_shape(const py::object& like, const std::vector<int64_t>& shape) {
  py::object torch = torch_module();
  return torch.attr("empty")(py::tuple(py::cast(shape)), py::arg("dtype") = torch.attr("float64"),
                             py::arg("device") = like.attr("device"));
}

const MeshD<3>& mesh3(const py::object& mesh) {
  if (!py::isinstance<MeshD<3>>(mesh))
    throw py::value_error("the device derivatives take a 3-d Mesh (Mesh[3])");
  const MeshD<3>& m = mesh.cast<const MeshD<3>&>();
  if (m.basis != Basis::Legendre || m.quadrature != Quadrature::GaussLobatto)
    throw py::value_error("only Legendre / GaussLobatto meshes have device kernels");
  return m;
}

// tensor [nc, Np] or [E, nc, Np] -> (E, nc, batched)
std::array<int64_t, 3> field_shape(const py::object& t, int64_t np, const char* name,
                                   int64_t comp_multiple = 1) {
  const std::vector<int64_t> s = shape_of(t);
  if (s.size() == 2 && s[1] == np && s[0] % comp_multiple == 0) return {1, s[0], 0};
  if (s.size() == 3 && s[2] == np && s[1] % comp_multiple == 0) return {s[0], s[1], 1};
  throw py::value_error(std::string(name) + " must be [nc, Np] or [E, nc, Np] with Np = " +
                        std::to_string(np));
}

void require_inv_jac(const py::object& inv, int64_t e, int64_t np, bool batched) {
  const std::vector<int64_t> s = shape_of(inv);
  const bool ok = batched ? (s.size() == 3 && s[0] == e && s[1] == 9 && s[2] == np)
                          : (s.size() == 2 && s[0] == 9 && s[1] == np);
  if (!ok) throw py::value_error("inv_jacobian must be [9, Np] (or [E, 9, Np] for a batch)");
}

// partial_derivative(tensor, mesh, inv_jacobian): [nc, Np] -> [3 nc, Np]
// (derivative index fastest, block 3 c + i), batched [E, ...] likewise.
py::object partial_derivative(const py::object& tensor, const py::object& mesh,
                              const py::object& inv_jacobian) {
  const MeshD<3>& m = mesh3(mesh);
  const int64_t np = m.number_of_grid_points();
  const auto [e, nc, batched] = field_shape(tensor, np, "tensor");
  require_inv_jac(inv_jacobian, e, np, batched);
  const double* u = static_cast<const double*>(device_ptr(tensor, "tensor"));
  const double* ij = static_cast<const double*>(device_ptr(inv_jacobian, "inv_jacobian"));
  std::vector<int64_t> out_shape = batched ? std::vector<int64_t>{e, 3 * nc, np}
                                           : std::vector<int64_t>{3 * nc, np};
  py::object out = empty_like_shape(tensor, out_shape);
  double* du = static_cast<double*>(device_ptr(out, "out"));
  void* stream = current_stream(tensor);
  const int32_t ext[3] = {m.ext[0], m.ext[1], m.ext[2]};
  check(sxd_partial_derivatives(static_cast<int>(e), ext, static_cast<int>(nc), u, ij, du,
                                stream));
  return out;
}

// logical_partial_derivative(tensor, mesh): [nc, Np] -> [3 nc, Np] with the
// logical derivative index fastest (the reference prepends the
// ElementLogical index).
py::object logical_partial_derivative(const py::object& tensor, const py::object& mesh) {
  const MeshD<3>& m = mesh3(mesh);
  const int64_t np = m.number_of_grid_points();
  const auto [e, nc, batched] = field_shape(tensor, np, "tensor");
  const double* u = static_cast<const double*>(device_ptr(tensor, "tensor"));
  // the C ABI writes direction-major [3][E][nc][Np]; reorder on the device
  py::object dirmajor = empty_like_shape(tensor, {3, e, nc, np});
  double* dl = static_cast<double*>(device_ptr(dirmajor, "out"));
  void* stream = current_stream(tensor);
  const int32_t ext[3] = {m.ext[0], m.ext[1], m.ext[2]};
  check(sxd_logical_partial_derivatives(static_cast<int>(e), ext, static_cast<int>(nc), u, dl,
                                        stream));
  py::object out = dirmajor.attr("permute")(1, 2, 0, 3).attr("contiguous")();
  std::vector<int64_t> out_shape = batched ? std::vector<int64_t>{e, 3 * nc, np}
                                           : std::vector<int64_t>{3 * nc, np};
  return out.attr("reshape")(py::tuple(py::cast(out_shape)));
}

// Transpose of partial_derivative for a fixed inverse Jacobian:
// [3 nc, Np] -> [nc, Np] (batched likewise).
py::object partial_derivative_adjoint(const py::object& grad, const py::object& mesh,
                                      const py::object& inv_jacobian) {
  const MeshD<3>& m = mesh3(mesh);
  const int64_t np = m.number_of_grid_points();
  const auto [e, nc3, batched] = field_shape(grad, np, "grad", 3);
  require_inv_jac(inv_jacobian, e, np, batched);
  const int64_t nc = nc3 / 3;
  const double* g = static_cast<const double*>(device_ptr(grad, "grad"));
  const double* ij = static_cast<const double*>(device_ptr(inv_jacobian, "inv_jacobian"));
  std::vector<int64_t> out_shape = batched ? std::vector<int64_t>{e, nc, np}
                                           : std::vector<int64_t>{nc, np};
  py::object out = empty_like_shape(grad, out_shape);
  double* ub = static_cast<double*>(device_ptr(out, "out"));
  void* stream = current_stream(grad);
  const int32_t ext[3] = {m.ext[0], m.ext[1], m.ext[2]};
  check(sxd_partial_derivatives_adjoint(static_cast<int>(e), ext, static_cast<int>(nc), g, ij,
                                        ub, stream));
  return out;
}

std::vector<double> collocation_points(const py::object& mesh) {
  if (!py::isinstance<MeshD<1>>(mesh))
    throw py::value_error("collocation_points takes a 1-d Mesh (Mesh[1])");
  const MeshD<1>& m = mesh.cast<const MeshD<1>&>();
  if (m.basis != Basis::Legendre || m.quadrature != Quadrature::GaussLobatto)
    throw py::value_error("only Legendre / GaussLobatto meshes are built");
  std::vector<double> x(m.ext[0]), w(m.ext[0]);
  check(sxd_lgl_points_and_weights(m.ext[0], x.data(), w.data()));
  return x;
}

std::vector<double> quadrature_weights(const py::object& mesh) {
  if (!py::isinstance<MeshD<1>>(mesh))
    throw py::value_error("quadrature_weights takes a 1-d Mesh (Mesh[1])");
  const MeshD<1>& m = mesh.cast<const MeshD<1>&>();
  std::vector<double> x(m.ext[0]), w(m.ext[0]);
  check(sxd_lgl_points_and_weights(m.ext[0], x.data(), w.data()));
  return w;
}

template <int Dim>
void bind_mesh(py::module_& mod, const char* name) {
  py::class_<MeshD<Dim>>(mod, name)
      .def(py::init([](int n, Basis b, Quadrature q) {
             MeshD<Dim> m;
             m.ext.fill(n);
             m.basis = b;
             m.quadrature = q;
             return m;
           }),
           py::arg("extents"), py::arg("basis"), py::arg("quadrature"))
      .def(py::init([](const std::vector<int>& ext, Basis b, Quadrature q) {
             if (ext.size() != Dim) throw py::value_error("wrong number of extents");
             MeshD<Dim> m;
             for (int d = 0; d < Dim; ++d) m.ext[d] = ext[d];
             m.basis = b;
             m.quadrature = q;
             return m;
           }),
           py::arg("extents"), py::arg("basis"), py::arg("quadrature"))
      .def_property_readonly("dim", [](const MeshD<Dim>&) { return Dim; })
      .def("extents", [](const MeshD<Dim>& m) { return std::vector<int>(m.ext.begin(), m.ext.end()); })
      .def("extents", [](const MeshD<Dim>& m, int d) { return m.ext.at(d); })
      .def("number_of_grid_points", &MeshD<Dim>::number_of_grid_points)
      .def("basis", [](const MeshD<Dim>& m) { return std::vector<Basis>(Dim, m.basis); })
      .def("quadrature",
           [](const MeshD<Dim>& m) { return std::vector<Quadrature>(Dim, m.quadrature); })
      .def("__repr__", [](const MeshD<Dim>& m) {
        std::string s = "Mesh" + std::to_string(Dim) + "D(";
        for (int d = 0; d < Dim; ++d) s += (d ? "," : "") + std::to_string(m.ext[d]);
        return s + ")";
      });
}

// ---- domain calls on a handle (sxd_domain*, as an integer) ---------------
sxd_domain* handle(uintptr_t h) {
  if (h == 0) throw py::value_error("null domain handle");
  return reinterpret_cast<sxd_domain*>(h);
}

#define SXD_PTR(t) static_cast<double*>(device_ptr(t, #t))
#define SXD_CPTR(t) static_cast<const double*>(device_ptr(t, #t))

}  // namespace

PYBIND11_MODULE(_Pybindings, m) {
  m.doc() = "spectre_amd device path: Spectral meshes, linear operators, DG right-hand "
            "sides and their adjoints (pybind11)";
  py::enum_<Basis>(m, "Basis")
      .value("Legendre", Basis::Legendre)
      .value("Chebyshev", Basis::Chebyshev)
      .value("FiniteDifference", Basis::FiniteDifference)
      .value("SphericalHarmonic", Basis::SphericalHarmonic);
  py::enum_<Quadrature>(m, "Quadrature")
      .value("GaussLobatto", Quadrature::GaussLobatto)
      .value("Gauss", Quadrature::Gauss)
      .value("CellCentered", Quadrature::CellCentered)
      .value("FaceCentered", Quadrature::FaceCentered);
  bind_mesh<1>(m, "Mesh1D");
  bind_mesh<2>(m, "Mesh2D");
  bind_mesh<3>(m, "Mesh3D");
  m.def("collocation_points", &collocation_points, py::arg("mesh"));
  m.def("quadrature_weights", &quadrature_weights, py::arg("mesh"));
  m.def("partial_derivative", &partial_derivative, py::arg("tensor"), py::arg("mesh"),
        py::arg("inv_jacobian"));
  m.def("logical_partial_derivative", &logical_partial_derivative, py::arg("tensor"),
        py::arg("mesh"));
  m.def("partial_derivative_adjoint", &partial_derivative_adjoint, py::arg("grad"),
        py::arg("mesh"), py::arg("inv_jacobian"));

  // whole-domain right-hand sides on a domain handle (dg.Domain.handle)
  m.def("scalar_wave_rhs", [](uintptr_t h, py::object u, py::object gamma2, py::object inv_jac,
                              py::object out) {
    const double *a = SXD_CPTR(u), *b = SXD_CPTR(gamma2), *c = SXD_CPTR(inv_jac);
    double* o = SXD_PTR(out);
    check(sxd_scalar_wave_rhs(handle(h), a, b, c, o, current_stream(u)));
    return out;
  });
  m.def("gh_rhs", [](uintptr_t h, py::object u, py::object gammas, py::object inv_jac,
                     py::object out) {
    const double *a = SXD_CPTR(u), *b = SXD_CPTR(gammas), *c = SXD_CPTR(inv_jac);
    double* o = SXD_PTR(out);
    check(sxd_gh_rhs(handle(h), a, b, c, o, current_stream(u)));
    return out;
  });
  m.def("scalar_wave_rhs_adjoint", [](uintptr_t h, py::object gbar, py::object gamma2,
                                      py::object inv_jac, py::object out) {
    const double *a = SXD_CPTR(gbar), *b = SXD_CPTR(gamma2), *c = SXD_CPTR(inv_jac);
    double* o = SXD_PTR(out);
    check(sxd_scalar_wave_rhs_adjoint(handle(h), a, b, c, o, current_stream(gbar)));
    return out;
  });
  m.def("gh_rhs_adjoint", [](uintptr_t h, py::object u, py::object gammas, py::object inv_jac,
                             py::object gbar, py::object out) {
    const double *a = SXD_CPTR(u), *b = SXD_CPTR(gammas), *c = SXD_CPTR(inv_jac),
                 *g = SXD_CPTR(gbar);
    double* o = SXD_PTR(out);
    check(sxd_gh_rhs_adjoint(handle(h), a, b, c, g, o, current_stream(u)));
    return out;
  });
  m.def("rk3_substep", [](int substep, double dt, py::object u0, py::object uk, py::object lk,
                          py::object out) {
    const int64_t n = out.attr("numel")().cast<int64_t>();
    const double *a = SXD_CPTR(u0), *b = SXD_CPTR(uk), *c = SXD_CPTR(lk);
    double* o = SXD_PTR(out);
    check(sxd_rk3_substep(substep, n, dt, a, b, c, o, current_stream(out)));
    return out;
  });
  m.def("rk3_adjoint_substep", [](int substep, double dt, py::object lam, py::object a,
                                  py::object mu, py::object lam_out) {
    const int64_t n = lam.attr("numel")().cast<int64_t>();
    double* lo = lam_out.is_none() ? nullptr : SXD_PTR(lam_out);
    const double *l = SXD_CPTR(lam), *x = SXD_CPTR(a);
    double* mp = SXD_PTR(mu);
    check(sxd_rk3_adjoint_substep(substep, n, dt, l, x, mp, lo, current_stream(lam)));
  });
}
