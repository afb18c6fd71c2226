// SparseSolvPy: Python module of the GPU solvers with the reference's names
// and keyword arguments (pybind/pybind.cpp:16-141), kept verbatim including
// its inconsistencies (SparseBuilder.getSize vs SparseBuilderC.get_size,
// SparseMat(builder, to_square=True)).  Vectors are numpy arrays; solve_*
// return (converged, x) with x a fresh array (x0 is taken by value).  The GIL
// is released while a solve runs on the GPU.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <SparseSolv/MatSolverICCG.hpp>
#include <SparseSolv/MatSolverMRTR.hpp>

namespace py = pybind11;
using namespace JPMRspace::SparseSolv;
using JPMRspace::dcomplex;

namespace {

template <typename T> using arr = py::array_t<T, py::array::c_style | py::array::forcecast>;

// b is read in place from the numpy buffer (a C-contiguous array of the
// value type is used as is; anything else is converted once by pybind11); the
// library uploads it straight from there.  x starts as a copy of x0 (the
// reference takes x0 by value) and the solve writes it in place.
template <typename Solver, typename Mat, typename T>
py::tuple solve_impl(Solver& self, double conv_cri, int max_ite, const Mat& A, arr<T> b, arr<T> x0) {
    const slv_int n = A.getSize();
    if (b.size() != n || x0.size() != n) throw std::invalid_argument("b and x0 must have A.getSize() entries");
    const T* bp = b.data();   // b stays referenced by this frame while the GIL is released
    py::array_t<T> x(n);
    std::copy(x0.data(), x0.data() + n, x.mutable_data());
    T* xp = x.mutable_data();
    bool ok;
    {
        py::gil_scoped_release nogil;
        // reference lambda: self.solve(conv_cri, max_ite, A, b, x0) on Eigen vectors (normB = b.norm())
        ok = self.solve(conv_cri, max_ite, norm_eigen_order(bp, n), A, bp, xp);
    }
    return py::make_tuple(ok, x);
}

}  // namespace

PYBIND11_MODULE(SparseSolvPy, m) {
    m.doc() = "SparseSolv bindings (MI355X GPU solvers)";

    py::class_<SparseBuilder>(m, "SparseBuilder")
        .def(py::init<slv_int>(), py::arg("size"))
        .def("getSize", &SparseBuilder::getSize)
        .def("initialize", py::overload_cast<slv_int>(&SparseBuilder::initialize), py::arg("size"))
        .def("reInitialize", &SparseBuilder::reInitialize, py::arg("size"))
        .def("add", &SparseBuilder::add, py::arg("row"), py::arg("col"), py::arg("val"))
        .def("resetMat", &SparseBuilder::resetMat);

    py::class_<SparseBuilderC>(m, "SparseBuilderC")
        .def(py::init<slv_int>(), py::arg("size"))
        .def("get_size", &SparseBuilderC::getSize)
        .def("initialize", py::overload_cast<slv_int>(&SparseBuilderC::initialize), py::arg("size"))
        .def("reInitialize", &SparseBuilderC::reInitialize, py::arg("size"))
        .def("add", &SparseBuilderC::add, py::arg("row"), py::arg("col"), py::arg("val"))
        .def("resetMat", &SparseBuilderC::resetMat);

    py::class_<SparseMat>(m, "SparseMat")
        .def(py::init<>())
        .def(py::init<SparseBuilder&, bool>(), py::arg("builder"), py::arg("to_square") = true)
        .def(py::init<slv_int, const std::vector<slv_int>&, const std::vector<slv_int>&, const std::vector<double>&>(),
             py::arg("n_nonzero"), py::arg("rows"), py::arg("cols"), py::arg("vals"))
        .def("getSize", &SparseMat::getSize)
        .def("getMaxCol", &SparseMat::getMaxCol)
        .def("print_mat", &SparseMat::printMat, py::arg("filename") = "Mat.csv");

    py::class_<SparseMatC>(m, "SparseMatC")
        .def(py::init<>())
        .def(py::init<SparseBuilderC&, bool>(), py::arg("builder"), py::arg("to_square") = true)
        .def(py::init<slv_int, const std::vector<slv_int>&, const std::vector<slv_int>&,
                      const std::vector<dcomplex>&>(),
             py::arg("n_nonzero"), py::arg("rows"), py::arg("cols"), py::arg("vals"))
        .def("getSize", &SparseMatC::getSize)
        .def("getMaxCol", &SparseMatC::getMaxCol)
        .def("print_mat", &SparseMatC::printMat, py::arg("filename") = "Mat.csv");

    py::class_<MatSolverBase>(m, "MatSolverBase")
        .def("setInitSolution", &MatSolverBase::setInitSolution)
        .def("setDiagScale", &MatSolverBase::setDiagScale)
        .def("setSaveBest", &MatSolverBase::setSaveBest)
        .def("setSaveLog", &MatSolverBase::setSaveLog)
        .def("setDirvegeType", &MatSolverBase::setDirvegeType)
        .def("setBadDivVal", &MatSolverBase::setBadDivVal)
        .def("setBadDivCount", &MatSolverBase::setBadDivCount)
        .def("setConvNormalizeType", &MatSolverBase::setConvNormalizeType)
        .def("setConvNormalizeConst", &MatSolverBase::setConvNormalizeConst)
        .def("getResidualLog", [](MatSolverBase& self) {
            std::vector<double> log;
            self.getResidualLog(log);
            return log;
        })
        // additions of the GPU library
        .def("setDevice", &MatSolverBase::setDevice)
        .def("setFactorPrecision", &MatSolverBase::setFactorPrecision)
        .def("setABMC", &MatSolverBase::setABMC, py::arg("block_size"))
        .def("getIterations", &MatSolverBase::getIterations)
        .def("getSetupMs", &MatSolverBase::getSetupMs)
        .def("getIterateMs", &MatSolverBase::getIterateMs);

    py::class_<MatSolverICCG, MatSolverBase>(m, "MatSolverICCG")
        .def(py::init<>())
        .def("setAccelFactor", &MatSolverICCG::setAccelFactor)
        .def("solve_real", &solve_impl<MatSolverICCG, SparseMat, double>, py::arg("conv_cri"), py::arg("max_ite"),
             py::arg("A"), py::arg("b"), py::arg("x0"))
        .def("solve_complex", &solve_impl<MatSolverICCG, SparseMatC, dcomplex>, py::arg("conv_cri"),
             py::arg("max_ite"), py::arg("A"), py::arg("b"), py::arg("x0"));

    py::class_<MatSolverMRTR, MatSolverBase>(m, "MatSolverMRTR")
        .def(py::init<>())
        .def("setAccelFactor", &MatSolverMRTR::setAccelFactor)
        .def("setPreconditionType", &MatSolverMRTR::setPreconditionType)
        .def("setDirectionFix", &MatSolverMRTR::setDirectionFix)
        .def("solve_real", &solve_impl<MatSolverMRTR, SparseMat, double>, py::arg("conv_cri"), py::arg("max_ite"),
             py::arg("A"), py::arg("b"), py::arg("x0"))
        .def("solve_complex", &solve_impl<MatSolverMRTR, SparseMatC, dcomplex>, py::arg("conv_cri"),
             py::arg("max_ite"), py::arg("A"), py::arg("b"), py::arg("x0"));
}
