// pybind11 module `py_module`: the Python surface of rdist_tensor's
// exafac.cpp_ext.py_module (exafac/cpp_ext/py_module.cpp:37-77), bound over
// the C-ABI of libexafac_hip.so (include/exafac_hip.h).
//
// Same class names, constructor signatures and methods as the reference.
// Additions are keyword-only extras (seed=, precision=) and accessors for the
// factors.  Differences by design:
//  * one process drives one GPU: Grid accepts only all-ones processor grids;
//  * parents are kept alive with py::keep_alive (the reference stores raw
//    references, grid.hpp:130, distmat.hpp:18-19, sampler.hpp:18, ...);
//  * the GIL is released around device work.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstdlib>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "exafac_hip.h"

namespace py = pybind11;

namespace {

void check(int rc) {
    if (rc == EX_OK) return;
    std::string msg = ex_last_error();
    if (rc == EX_ERR_ARG) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

ex_ctx* context() {
    static ex_ctx* ctx = nullptr;
    if (!ctx) {
        int dev = 0;
        if (const char* e = std::getenv("EXAFAC_DEVICE")) dev = std::atoi(e);
        check(ex_ctx_create(dev, &ctx));
    }
    return ctx;
}

int precision_code(const std::string& p) {
    if (p == "fp64") return EX_FP64;
    if (p == "mixed") return EX_MIXED;
    throw py::value_error("precision must be 'fp64' or 'mixed'");
}

// ---------------------------------------------------------------- grids
// Grid (grid.hpp:16-122): one rank, so every processor-grid extent is 1.
struct Grid {
    std::vector<int> proc_dims;
    explicit Grid(py::array_t<int> dims) {
        auto d = dims.unchecked<1>();
        for (py::ssize_t i = 0; i < d.shape(0); i++) {
            if (d(i) != 1)
                throw std::runtime_error("this build drives one GPU as a single rank: every processor-grid "
                                         "dimension must be 1");
            proc_dims.push_back(d(i));
        }
        if (proc_dims.empty()) throw py::value_error("processor grid must have at least one dimension");
    }
    int get_dimension() const { return (int)proc_dims.size(); }
};

// TensorGrid (grid.hpp:124-174).  At one rank the padded row count of mode i
// is round_to_nearest_integer(I_i, 1) = I_i + 1 (common.h:196-198).
struct TensorGrid {
    std::vector<int64_t> tensor_dims;
    Grid* grid;
    TensorGrid(py::array_t<int> dims, Grid& g) : grid(&g) {
        auto d = dims.unchecked<1>();
        for (py::ssize_t i = 0; i < d.shape(0); i++) {
            if (d(i) <= 0) throw py::value_error("tensor dimensions must be positive");
            tensor_dims.push_back(d(i));
        }
        if ((int)tensor_dims.size() != g.get_dimension())
            throw std::runtime_error("tensor grid dimension must match the processor grid dimension");
    }
    int dim() const { return (int)tensor_dims.size(); }
};

// DistMat1D (distmat.hpp:15-193): a row-block factor matrix; at one rank the
// whole (I+1) x cols matrix (the last row is the padding row).
struct DistMat1D {
    uint64_t cols, slice_dim, rows, padded_rows;
    py::array_t<double> data;
    DistMat1D(uint64_t c, TensorGrid& tg, uint64_t sd) : cols(c), slice_dim(sd) {
        if (sd >= tg.tensor_dims.size()) throw py::value_error("slice_dim out of range");
        rows = (uint64_t)tg.tensor_dims[sd];
        padded_rows = rows + 1;
        data = py::array_t<double>({(py::ssize_t)padded_rows, (py::ssize_t)cols});
        std::fill(data.mutable_data(), data.mutable_data() + padded_rows * cols, 0.0);
    }
};

// ---------------------------------------------------------------- low-rank tensor
struct LowRankTensor {
    uint64_t rank;
    TensorGrid* tg;
    ex_factors* f = nullptr;
    LowRankTensor(uint64_t R, TensorGrid& g) : rank(R), tg(&g) {
        if (R == 0) throw py::value_error("rank must be positive");
        check(ex_factors_create(context(), g.dim(), g.tensor_dims.data(), (int64_t)R, &f));
    }
    ~LowRankTensor() {
        if (f) ex_factors_destroy(f);
    }
    int dim() const { return tg->dim(); }
    // low_rank_tensor.hpp:27-33
    void initialize_factors_deterministic() {
        py::gil_scoped_release nogil;
        check(ex_factors_init_deterministic(f));
    }
    // low_rank_tensor.hpp:35-41: N(0,1) entries, columns renormalised, sigma = 1.
    // The draws come from numpy.random.default_rng(seed), mode by mode.
    void initialize_factors_gaussian_random(py::object seed) {
        py::object rng = py::module_::import("numpy.random").attr("default_rng")(seed);
        for (int i = 0; i < dim(); i++) {
            py::array_t<double, py::array::c_style | py::array::forcecast> a =
                rng.attr("standard_normal")(py::make_tuple(tg->tensor_dims[i], rank));
            check(ex_factors_set(f, i, a.data()));
            check(ex_factors_renormalize(f, i, nullptr));
        }
        std::vector<double> ones(rank, 1.0);
        check(ex_factors_set_sigma(f, ones.data()));
    }
    // low_rank_tensor.hpp:44-54: cosine-initialise every factor (without
    // renormalising) and print its Gram matrix.
    void test_gram_matrix_computation() {
        check(ex_factors_init_deterministic(f));
        // the reference prints the Gram of the un-normalised cosine factor
        for (int i = 0; i < dim(); i++) {
            int64_t I = tg->tensor_dims[i];
            std::vector<double> U((size_t)I * rank);
            for (int64_t r = 0; r < I; r++)
                for (uint64_t c = 0; c < rank; c++) U[r * rank + c] = std::cos((double)(uint64_t)(r * rank + c));
            check(ex_factors_set(f, i, U.data()));
            std::vector<double> G(rank * rank);
            check(ex_factors_gram(f, i, G.data()));
            py::print("Gram matrix for factor", i);
            py::print("------------------------");
            for (uint64_t a = 0; a < rank; a++) {
                py::list row;
                for (uint64_t b = 0; b < rank; b++) row.append(G[a * rank + b]);
                py::print(row);
            }
            py::print("------------------------");
        }
    }
    py::array_t<double> get_factor(int i) {
        if (i < 0 || i >= dim()) throw py::index_error("mode out of range");
        py::array_t<double> out({(py::ssize_t)tg->tensor_dims[i], (py::ssize_t)rank});
        check(ex_factors_get(f, i, out.mutable_data()));
        return out;
    }
    void set_factor(int i, py::array_t<double, py::array::c_style | py::array::forcecast> a) {
        if (i < 0 || i >= dim()) throw py::index_error("mode out of range");
        if (a.ndim() != 2 || a.shape(0) != tg->tensor_dims[i] || (uint64_t)a.shape(1) != rank)
            throw py::value_error("factor must have shape (I_mode, R)");
        check(ex_factors_set(f, i, a.data()));
    }
    py::array_t<double> get_sigma() {
        py::array_t<double> out((py::ssize_t)rank);
        check(ex_factors_get_sigma(f, out.mutable_data()));
        return out;
    }
    void set_sigma(py::array_t<double, py::array::c_style | py::array::forcecast> a) {
        if (a.ndim() != 1 || (uint64_t)a.shape(0) != rank) throw py::value_error("sigma must have shape (R,)");
        check(ex_factors_set_sigma(f, a.data()));
    }
};

// ---------------------------------------------------------------- sparse tensor
struct SparseTensor {
    TensorGrid* tg;
    ex_tensor* t = nullptr;
    std::string preprocessing;
    int64_t nnz = 0;
    // sparse_tensor.hpp:36-96
    SparseTensor(TensorGrid& g, py::array_t<uint32_t, py::array::c_style | py::array::forcecast> idx,
                 py::array_t<double, py::array::c_style | py::array::forcecast> vals, std::string pre)
        : tg(&g), preprocessing(pre) {
        const int N = g.dim();
        if (idx.ndim() != 2 || idx.shape(1) != N) throw py::value_error("indices must have shape (nnz, N)");
        if (vals.ndim() != 1 || vals.shape(0) != idx.shape(0)) throw py::value_error("values must have shape (nnz,)");
        int kind = EX_PRE_NONE;
        if (pre == "log_count") kind = EX_PRE_LOG_COUNT;
        else if (pre == "ones") kind = EX_PRE_ONES;
        nnz = idx.shape(0);
        const uint32_t* ip = idx.data();
        const double* vp = vals.data();
        py::gil_scoped_release nogil;
        check(ex_tensor_create(context(), N, g.tensor_dims.data(), nnz, ip, vp, kind, &t));
    }
    // sparse_tensor.hpp:102-215: each row of each of the 3 modes joins one of
    // Q blocks uniformly at random; every (u, v, w) whose rows share a block is
    // a nonzero of value 1.  Draws from std::mt19937_64(seed).
    SparseTensor(TensorGrid& g, uint64_t I, uint64_t N, uint64_t Q, uint64_t seed) : tg(&g), preprocessing("none") {
        if (N != 3) throw py::value_error("the synthetic block tensor is defined for N = 3 only");
        if (Q == 0 || I == 0) throw py::value_error("I and Q must be positive");
        if (g.dim() != 3) throw py::value_error("tensor grid must have 3 modes");
        for (int i = 0; i < 3; i++)
            if ((uint64_t)g.tensor_dims[i] < I) throw py::value_error("tensor grid dimensions must be at least I");
        std::mt19937_64 gen(seed);
        std::uniform_int_distribution<uint64_t> dist(0, Q - 1);
        std::vector<std::vector<std::vector<uint32_t>>> bl(3, std::vector<std::vector<uint32_t>>(Q));
        for (int i = 0; i < 3; i++)
            for (uint64_t k = 0; k < I; k++) bl[i][dist(gen)].push_back((uint32_t)k);
        std::vector<uint32_t> idx;
        for (uint64_t j = 0; j < Q; j++)
            for (uint32_t u : bl[0][j])
                for (uint32_t v : bl[1][j])
                    for (uint32_t w : bl[2][j]) {
                        idx.push_back(u);
                        idx.push_back(v);
                        idx.push_back(w);
                    }
        nnz = (int64_t)(idx.size() / 3);
        std::vector<double> vals((size_t)nnz, 1.0);
        py::gil_scoped_release nogil;
        check(ex_tensor_create(context(), 3, g.tensor_dims.data(), nnz, idx.data(), vals.data(), EX_PRE_NONE, &t));
    }
    ~SparseTensor() {
        if (t) ex_tensor_destroy(t);
    }
    double norm() {
        double v = 0;
        check(ex_tensor_norm(t, &v));
        return v;
    }
};

// ---------------------------------------------------------------- samplers
struct Sampler {
    ex_sampler* s = nullptr;
    LowRankTensor* lrt = nullptr;
    std::string precision = "fp64";
    virtual ~Sampler() {
        if (s) ex_sampler_destroy(s);
    }
    // Sampler::KRPDrawSamples (sampler.hpp:34-38) -> (samples (J, N) uint32, weights (J,))
    py::tuple draw(int64_t J, int mode, uint64_t sweep) {
        const int N = lrt->dim();
        py::array_t<uint32_t> samples({(py::ssize_t)J, (py::ssize_t)N});
        py::array_t<double> weights((py::ssize_t)J);
        uint32_t* sp = samples.mutable_data();
        double* wp = weights.mutable_data();
        {
            py::gil_scoped_release nogil;
            check(ex_sampler_draw(s, J, mode, sweep, nullptr, nullptr, sp, wp));
        }
        return py::make_tuple(samples, weights);
    }
    void update_sampler(int mode) {
        py::gil_scoped_release nogil;
        check(ex_sampler_update(s, mode));
    }
};

// cp_arls_lev.hpp:14-23
struct CP_ARLS_LEV : Sampler {
    CP_ARLS_LEV(LowRankTensor& t, uint64_t seed, std::string prec) {
        lrt = &t;
        precision = prec;
        int pc = precision_code(prec);
        py::gil_scoped_release nogil;
        check(ex_sampler_create(context(), t.f, EX_SAMPLER_CP_ARLS_LEV, pc, seed, &s));
    }
};

// sts_cp.hpp:32-57
struct STS_CP : Sampler {
    STS_CP(LowRankTensor& t, uint64_t seed, std::string prec) {
        lrt = &t;
        precision = prec;
        int pc = precision_code(prec);
        py::gil_scoped_release nogil;
        check(ex_sampler_create(context(), t.f, EX_SAMPLER_STS_CP, pc, seed, &s));
    }
};

// ---------------------------------------------------------------- optimizers
struct ALS_Optimizer {
    ex_als* a = nullptr;
    virtual ~ALS_Optimizer() {
        if (a) ex_als_destroy(a);
    }
    void initialize_ground_truth_for_als() {
        if (!a) throw std::runtime_error("optimizer not constructed");
        py::gil_scoped_release nogil;
        check(ex_als_init_ground_truth(a));
    }
    void execute_ALS_rounds(uint64_t rounds, uint64_t J, uint32_t epoch) {
        py::gil_scoped_release nogil;
        check(ex_als_run(a, (int64_t)rounds, (int64_t)J, (int64_t)epoch));
    }
    std::string get_statistics_json() {
        size_t len = 0;
        check(ex_als_stats_json(a, nullptr, &len));
        std::string s(len, '\0');
        check(ex_als_stats_json(a, s.data(), &len));
        s.resize(len > 0 ? len - 1 : 0);
        return s;
    }
    double compute_exact_fit() {
        double fit = 0;
        py::gil_scoped_release nogil;
        check(ex_als_fit(a, &fit));
        return fit;
    }
    void execute_ALS_step(int mode, uint64_t J, uint64_t sweep) {
        py::gil_scoped_release nogil;
        check(ex_als_step(a, mode, (int64_t)J, sweep));
    }
    void deinitialize() {}
};

struct AccumulatorStationary : ALS_Optimizer {
    AccumulatorStationary(SparseTensor& t, LowRankTensor& f, Sampler& s) {
        check(ex_als_create(context(), t.t, f.f, s.s, EX_DIST_ACCUMULATOR_STATIONARY, precision_code(s.precision), &a));
    }
};

struct TensorStationary : ALS_Optimizer {
    TensorStationary(SparseTensor& t, LowRankTensor& f, Sampler& s) {
        check(ex_als_create(context(), t.t, f.f, s.s, EX_DIST_TENSOR_STATIONARY, precision_code(s.precision), &a));
    }
};

// exact_als.hpp:16-51; `precision` ("fp64" or "mixed") has no reference
// counterpart.  execute_ALS_step / execute_ALS_rounds ignore J.
struct ExactALS : ALS_Optimizer {
    ExactALS(SparseTensor& t, LowRankTensor& f, const std::string& prec) {
        check(ex_exact_als_create(context(), t.t, f.f, precision_code(prec), &a));
    }
};

}  // namespace

PYBIND11_MODULE(py_module, m) {
    m.doc() = "rdist_tensor's exafac.cpp_ext.py_module surface on one MI355X (HIP kernels behind a C-ABI)";
    py::class_<Grid>(m, "Grid")
        .def(py::init<py::array_t<int>>())
        .def("get_dimension", &Grid::get_dimension);
    py::class_<TensorGrid>(m, "TensorGrid")
        .def(py::init<py::array_t<int>, Grid&>(), py::keep_alive<1, 3>())
        .def_property_readonly("tensor_dims", [](const TensorGrid& t) { return t.tensor_dims; });
    py::class_<DistMat1D>(m, "DistMat1D")
        .def(py::init<uint64_t, TensorGrid&, uint64_t>(), py::keep_alive<1, 3>())
        .def_readonly("rows", &DistMat1D::rows)
        .def_readonly("cols", &DistMat1D::cols)
        .def_readonly("padded_rows", &DistMat1D::padded_rows)
        .def_readonly("data", &DistMat1D::data);
    py::class_<LowRankTensor>(m, "LowRankTensor")
        .def(py::init<uint64_t, TensorGrid&>(), py::keep_alive<1, 3>())
        .def("test_gram_matrix_computation", &LowRankTensor::test_gram_matrix_computation)
        .def("initialize_factors_deterministic", &LowRankTensor::initialize_factors_deterministic)
        .def("initialize_factors_gaussian_random", &LowRankTensor::initialize_factors_gaussian_random,
             py::arg("seed") = py::none())
        .def("get_factor", &LowRankTensor::get_factor)
        .def("set_factor", &LowRankTensor::set_factor)
        .def("get_sigma", &LowRankTensor::get_sigma)
        .def("set_sigma", &LowRankTensor::set_sigma)
        .def_readonly("rank", &LowRankTensor::rank)
        .def_property_readonly("dim", &LowRankTensor::dim);
    py::class_<SparseTensor>(m, "SparseTensor")
        .def(py::init<TensorGrid&, py::array_t<uint32_t, py::array::c_style | py::array::forcecast>,
                      py::array_t<double, py::array::c_style | py::array::forcecast>, std::string>(),
             py::keep_alive<1, 2>())
        .def(py::init<TensorGrid&, uint64_t, uint64_t, uint64_t, uint64_t>(), py::arg("tensor_grid"), py::arg("I"),
             py::arg("N"), py::arg("Q"), py::arg("seed") = 0, py::keep_alive<1, 2>())
        .def_property_readonly("tensor_norm", &SparseTensor::norm)
        .def_readonly("nnz", &SparseTensor::nnz)
        .def_readonly("preprocessing", &SparseTensor::preprocessing);
    py::class_<ALS_Optimizer>(m, "ALS_Optimizer")
        .def("initialize_ground_truth_for_als", &ALS_Optimizer::initialize_ground_truth_for_als)
        .def("execute_ALS_rounds", &ALS_Optimizer::execute_ALS_rounds)
        .def("get_statistics_json", &ALS_Optimizer::get_statistics_json)
        .def("compute_exact_fit", &ALS_Optimizer::compute_exact_fit)
        .def("deinitialize", &ALS_Optimizer::deinitialize)
        .def("execute_ALS_step", &ALS_Optimizer::execute_ALS_step, py::arg("mode"), py::arg("J"),
             py::arg("sweep") = 1);
    py::class_<ExactALS, ALS_Optimizer>(m, "ExactALS")
        .def(py::init<SparseTensor&, LowRankTensor&, const std::string&>(), py::arg("ground_truth"),
             py::arg("low_rank_tensor"), py::arg("precision") = "fp64", py::keep_alive<1, 2>(),
             py::keep_alive<1, 3>());
    py::class_<TensorStationary, ALS_Optimizer>(m, "TensorStationary")
        .def(py::init<SparseTensor&, LowRankTensor&, Sampler&>(), py::keep_alive<1, 2>(), py::keep_alive<1, 3>(),
             py::keep_alive<1, 4>());
    py::class_<AccumulatorStationary, ALS_Optimizer>(m, "AccumulatorStationary")
        .def(py::init<SparseTensor&, LowRankTensor&, Sampler&>(), py::keep_alive<1, 2>(), py::keep_alive<1, 3>(),
             py::keep_alive<1, 4>());
    py::class_<Sampler>(m, "Sampler")
        .def("draw", &Sampler::draw, py::arg("J"), py::arg("mode"), py::arg("sweep") = 1)
        .def("update_sampler", &Sampler::update_sampler)
        .def_readonly("precision", &Sampler::precision);
    py::class_<CP_ARLS_LEV, Sampler>(m, "CP_ARLS_LEV")
        .def(py::init<LowRankTensor&, uint64_t, std::string>(), py::arg("low_rank_tensor"), py::arg("seed") = 0,
             py::arg("precision") = "fp64", py::keep_alive<1, 2>());
    py::class_<STS_CP, Sampler>(m, "STS_CP")
        .def(py::init<LowRankTensor&, uint64_t, std::string>(), py::arg("low_rank_tensor"), py::arg("seed") = 0,
             py::arg("precision") = "fp64", py::keep_alive<1, 2>());
}
