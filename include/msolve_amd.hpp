// msolve_amd.hpp -- header-only C++17 host layer over the C ABI
// (msolve_amd.h), mirroring the reference's object model and plug-in
// interfaces for this path so C++ callers can use it the way MSolve's C#
// code uses its own classes:
//
//   msolve::Model                 ISAAR.MSolve.PreProcessor/Model.cs
//                                 (nodes, Hexa8 elements, constraints, loads,
//                                 ConnectDataStructures = EnumerateDOFs +
//                                 AssignNodalLoads, Model.cs:165-238,289-296)
//   msolve::Hexa8MeshGenerator    SimpleMeshGenerator/Hexa8MeshGenerator.cs
//                                 (node/element numbering :508-568,1080-1089)
//   msolve::ElementStructuralStiffnessProvider
//                                 Providers/ElementStructuralStiffnessProvider.cs
//                                 (IElementMatrixProvider.Matrix, batched)
//   msolve::SolverPCG             Solvers/PCGSkyline/SolverPCGSkyline.cs
//                                 (Initialize / Solve(maxIterations, tolerance)
//                                 / CurrentIteration / VectorX)
//   msolve::MonteCarloAnalyzerWithStochasticMaterial
//                                 Analyzers/MonteCarloAnalyzerWithStochasticMaterial.cs
//                                 (Initialize / Solve over all samples at once)
//
// Errors surface as msolve::MsaError (the reference throws
// InvalidOperationException / ArgumentException).
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "msolve_amd.h"

namespace msolve {

class MsaError : public std::runtime_error {
  public:
    MsaError(int rc, const std::string &msg)
        : std::runtime_error("msolve_amd error " + std::to_string(rc) + ": " + msg), rc(rc) {}
    int rc;
};

inline void check(int rc)
{
    if (rc != MSA_OK) throw MsaError(rc, msa_last_error());
}

enum class DOFType { X = 0, Y = 1, Z = 2 };

struct Load {
    int node;
    DOFType dof;
    double amount;
};

class Context {
  public:
    explicit Context(int device = 0) { check(msa_ctx_create(device, &h_)); }
    ~Context() { msa_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    msa_ctx *get() const { return h_; }

  private:
    msa_ctx *h_ = nullptr;
};

// Model: node coordinates, Hexa8 connectivity (0-based), per-DOF constraints
// and nodal loads.  ConnectDataStructures numbers the free DOFs (nodes
// ascending, X/Y/Z, constrained -> -1) and assigns the loads with `=`
// (a second load on the same DOF overwrites the first, Model.cs:236).
class Model {
  public:
    std::vector<double> xyz;            // [n_nodes][3]
    std::vector<int> conn;              // [n_elem][8]
    std::vector<unsigned char> constrained;  // [n_nodes][3]
    std::vector<Load> Loads;
    std::vector<int> NodalDOFsDictionary;    // [n_nodes][3], -1 = constrained
    std::vector<double> Forces;              // [TotalDOFs]
    int TotalDOFs = 0;
    double domain[3] = {1.0, 1.0, 1.0};

    int NodeCount() const { return (int)(xyz.size() / 3); }
    int ElementCount() const { return (int)(conn.size() / 8); }

    void ConnectDataStructures()
    {
        constrained.resize(xyz.size(), 0);
        NodalDOFsDictionary.assign(xyz.size(), -1);
        check(msa_enumerate_dofs(NodeCount(), constrained.data(), NodalDOFsDictionary.data(),
                                 &TotalDOFs));
        Forces.assign(TotalDOFs, 0.0);
        for (const Load &l : Loads) {
            const int d = NodalDOFsDictionary[3 * l.node + (int)l.dof];
            if (d >= 0) Forces[d] = l.amount;
        }
    }
};

struct Hexa8MeshGenerator {
    // Box [0,lx]x[0,ly]x[0,lz] with nex x ney x nez elements; coordinates
    // accumulated by repeated addition as the reference generator does.
    static Model Box(int nex, int ney, int nez, double lx = 1, double ly = 1, double lz = 1)
    {
        Model m;
        const int nX = nex + 1, nY = ney + 1, nZ = nez + 1;
        m.xyz.resize(3 * (size_t)nX * nY * nZ);
        double p[3] = {0, 0, 0};
        size_t node = 0;
        for (int i = 0; i < nZ; i++) {
            for (int j = 0; j < nY; j++) {
                for (int k = 0; k < nX; k++) {
                    m.xyz[3 * node] = p[0];
                    m.xyz[3 * node + 1] = p[1];
                    m.xyz[3 * node + 2] = p[2];
                    node++;
                    p[0] += lx / nex;
                }
                p[0] = 0;
                p[1] += ly / ney;
            }
            p[1] = 0;
            p[2] += lz / nez;
        }
        for (int i = 0; i < nez; i++)
            for (int j = 0; j < ney; j++)
                for (int k = 0; k < nex; k++) {
                    const int b = nX * nY * i + nX * j + k, t = b + nX * nY;
                    const int c[8] = {b, b + 1, b + nX + 1, b + nX, t, t + 1, t + nX + 1, t + nX};
                    m.conn.insert(m.conn.end(), c, c + 8);
                }
        m.constrained.assign(m.xyz.size(), 0);
        m.domain[0] = lx;
        m.domain[1] = ly;
        m.domain[2] = lz;
        return m;
    }

    // The workload's cantilever: x = 0 clamped, total load in Z spread over
    // the x = lx face.
    static Model Cantilever(int n, double total_load = -100.0)
    {
        Model m = Box(n, n, n);
        const int nX = n + 1;
        std::vector<int> face;
        for (int node = 0; node < m.NodeCount(); node++) {
            if (node % nX == 0)
                for (int d = 0; d < 3; d++) m.constrained[3 * node + d] = 1;
            if (node % nX == n) face.push_back(node);
        }
        for (int node : face) m.Loads.push_back({node, DOFType::Z, total_load / face.size()});
        m.ConnectDataStructures();
        return m;
    }
};

class Mesh {
  public:
    Mesh(Context &ctx, const Model &m)
    {
        check(msa_mesh_create(ctx.get(), m.NodeCount(), m.xyz.data(), m.ElementCount(),
                              m.conn.data(), m.NodalDOFsDictionary.data(), m.TotalDOFs, &h_));
    }
    ~Mesh() { msa_mesh_destroy(h_); }
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;
    msa_mesh *get() const { return h_; }

  private:
    msa_mesh *h_ = nullptr;
};

// IElementMatrixProvider.Matrix for a whole element list at once: returns
// the packed SymmetricMatrix2D data (300 per element).
class ElementStructuralStiffnessProvider {
  public:
    explicit ElementStructuralStiffnessProvider(Context &ctx) : ctx_(ctx) {}
    std::vector<double> Matrix(const Model &m, const std::vector<double> &E, double nu)
    {
        const int ne = m.ElementCount();
        std::vector<double> xyz8(24 * (size_t)ne), ke(300 * (size_t)ne);
        std::vector<int> st(ne);
        for (int e = 0; e < ne; e++)
            for (int a = 0; a < 8; a++)
                for (int d = 0; d < 3; d++)
                    xyz8[24 * (size_t)e + 3 * a + d] = m.xyz[3 * (size_t)m.conn[8 * (size_t)e + a] + d];
        check(msa_ke_packed(ctx_.get(), ne, xyz8.data(), E.data(), nu, ke.data(), st.data()));
        for (int e = 0; e < ne; e++)
            if (st[e] != MSA_OK)
                throw MsaError(st[e], "element " + std::to_string(e) + ": Jacobian determinant below 1e-8");
        return ke;
    }

  private:
    Context &ctx_;
};

class Batch {
  public:
    Batch(Context &ctx, const std::vector<Mesh *> &meshes, const std::vector<int> &samples,
          int precision, int tile = 0)
    {
        std::vector<msa_mesh *> h;
        for (Mesh *m : meshes) h.push_back(m->get());
        check(msa_batch_create(ctx.get(), (int)h.size(), h.data(), samples.data(), precision, tile,
                               &h_));
    }
    ~Batch() { msa_batch_destroy(h_); }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    msa_batch *get() const { return h_; }

  private:
    msa_batch *h_ = nullptr;
};

// SolverPCG<SkylineMatrix2D<double>> surface for one model and one modulus
// field: Initialize() assembles, Solve() runs Jacobi-PCG on the GPU.
class SolverPCG {
  public:
    SolverPCG(Context &ctx, const Model &m, std::vector<double> E, double nu = 0.3)
        : ctx_(ctx), model_(m), E_(std::move(E)), nu_(nu) {}

    // MSA_PRECOND_JACOBI (the reference's choice) or MSA_PRECOND_MULTIGRID;
    // set before Initialize()
    int Preconditioner = MSA_PRECOND_JACOBI;

    void Initialize()
    {
        mesh_.reset(new Mesh(ctx_, model_));
        batch_.reset(new Batch(ctx_, {mesh_.get()}, {1}, MSA_FP64, 1));
        if (Preconditioner != MSA_PRECOND_JACOBI)
            check(msa_batch_set_preconditioner(batch_->get(), Preconditioner, 0, 0.0));
        check(msa_batch_set_modulus(batch_->get(), 0, E_.data()));
        check(msa_batch_set_rhs(batch_->get(), 0, model_.Forces.data()));
        check(msa_batch_assemble(batch_->get(), nu_));
    }

    // Solve() of SolverPCGSkyline.cs:102-108 uses 20000 iterations and 1e-7.
    void Solve(int maxIterations = 20000, double tolerance = 1e-7, int graphIters = 16)
    {
        if (!batch_) Initialize();
        check(msa_batch_solve(batch_->get(), tolerance, maxIterations, 0.5, 1, graphIters));
        int it = 0, outer = 0, status = 0;
        double rel = 0;
        check(msa_batch_get_stats(batch_->get(), 0, &it, &outer, &rel, &status));
        VectorX.assign(model_.TotalDOFs, 0.0);
        check(msa_batch_get_solution(batch_->get(), 0, 0, VectorX.data()));
        Status = status;
        Products = it;
        RelativeResidual = rel;
        // loop index at the break: one less than the products on convergence
        CurrentIteration = (status == MSA_OK && it > 0) ? it - 1 : it;
    }

    std::vector<double> VectorX;
    int CurrentIteration = 0, Products = 0, Status = -1;
    double RelativeResidual = 0;

  private:
    Context &ctx_;
    const Model &model_;
    std::vector<double> E_;
    double nu_;
    std::unique_ptr<Mesh> mesh_;
    std::unique_ptr<Batch> batch_;
};

// SolverSkyline (SolverSkyline.cs:55-205) for one model and one modulus
// field on the GPU band Cholesky: Initialize() assembles, Factorize() is
// SkylineMatrix2D.Factorize(tolerance, zems, zemCols) (SkylineMatrix2D.cs:
// 63-158: a pivot under tolerance is a zero-energy mode; tolerance 0 makes
// a non-positive pivot an error instead), Solve() the forward/backward
// substitution into VectorX.
class SolverSkyline {
  public:
    SolverSkyline(Context &ctx, const Model &m, std::vector<double> E, double nu = 0.3)
        : ctx_(ctx), model_(m), E_(std::move(E)), nu_(nu) {}

    void Initialize()
    {
        mesh_.reset(new Mesh(ctx_, model_));
        batch_.reset(new Batch(ctx_, {mesh_.get()}, {1}, MSA_FP64, 1));
        check(msa_batch_set_solver(batch_->get(), MSA_SOLVER_DIRECT));
        check(msa_batch_set_modulus(batch_->get(), 0, E_.data()));
        check(msa_batch_set_rhs(batch_->get(), 0, model_.Forces.data()));
        check(msa_batch_assemble(batch_->get(), nu_));
    }

    void Factorize(double tolerance, std::vector<std::vector<double>> &zems,
                   std::vector<int> &zemCols, int maxModes = 64)
    {
        if (!batch_) Initialize();
        check(msa_batch_set_zem(batch_->get(), tolerance, tolerance > 0 ? maxModes : 0));
        check(msa_batch_factorize(batch_->get()));
        zems.clear();
        zemCols.clear();
        if (tolerance <= 0) return;
        int n = 0;
        std::vector<int> cols(maxModes);
        std::vector<double> modes((size_t)maxModes * model_.TotalDOFs);
        check(msa_batch_get_zem(batch_->get(), 0, 0, maxModes, &n, cols.data(), modes.data()));
        for (int k = 0; k < n; k++) {
            zemCols.push_back(cols[k]);
            zems.emplace_back(modes.begin() + (size_t)k * model_.TotalDOFs,
                              modes.begin() + (size_t)(k + 1) * model_.TotalDOFs);
        }
    }

    void Solve()
    {
        if (!batch_) Initialize();
        check(msa_batch_direct_solve(batch_->get()));
        int status = 0;
        check(msa_batch_get_stats(batch_->get(), 0, nullptr, nullptr, nullptr, &status));
        VectorX.assign(model_.TotalDOFs, 0.0);
        check(msa_batch_get_solution(batch_->get(), 0, 0, VectorX.data()));
        Status = status;
    }

    std::vector<double> VectorX;
    int Status = -1;

  private:
    Context &ctx_;
    const Model &model_;
    std::vector<double> E_;
    double nu_;
    std::unique_ptr<Mesh> mesh_;
    std::unique_ptr<Batch> batch_;
};

// Karhunen-Loeve table in the fields of the reference's KL file
// (LognormalFileStochasticCoefficientsProvider.cs:90-109).
struct KLTable {
    int M = 0, kmax = 0;
    std::vector<double> w1d, a1d;  // [3][kmax]
    std::vector<int> evpPos;       // [M][3], 1-based
    std::vector<double> eigs;      // [M]
};

// All samples of all mesh classes solved as one device job.
class MonteCarloAnalyzerWithStochasticMaterial {
  public:
    MonteCarloAnalyzerWithStochasticMaterial(Context &ctx, std::vector<const Model *> models,
                                             std::vector<int> samples, KLTable table,
                                             std::vector<double> randomNumbers, double meanE,
                                             double cov, double nu, int precision)
        : ctx_(ctx), models_(std::move(models)), samples_(std::move(samples)),
          table_(std::move(table)), xi_(std::move(randomNumbers)), meanE_(meanE), cov_(cov),
          nu_(nu), precision_(precision) {}

    int Preconditioner = MSA_PRECOND_JACOBI;  // set before Initialize()

    void Initialize()
    {
        std::vector<Mesh *> ptrs;
        for (const Model *m : models_) {
            meshes_.emplace_back(new Mesh(ctx_, *m));
            ptrs.push_back(meshes_.back().get());
        }
        batch_.reset(new Batch(ctx_, ptrs, samples_, precision_, 0));
        if (Preconditioner != MSA_PRECOND_JACOBI)
            check(msa_batch_set_preconditioner(batch_->get(), Preconditioner, 0, 0.0));
        size_t off = 0;
        for (size_t c = 0; c < models_.size(); c++) {
            check(msa_batch_set_kl(batch_->get(), (int)c, table_.M, table_.kmax, table_.w1d.data(),
                                   table_.a1d.data(), table_.evpPos.data(), table_.eigs.data(),
                                   models_[c]->domain, meanE_, cov_));
            check(msa_batch_set_xi(batch_->get(), (int)c, xi_.data() + off * table_.M));
            check(msa_batch_set_rhs(batch_->get(), (int)c, models_[c]->Forces.data()));
            off += samples_[c];
        }
    }

    void Solve(double tolerance = 1e-10, int maxIterations = 20000, double innerTol = 1e-4,
               int maxOuter = 30, int graphIters = 16)
    {
        if (!batch_) Initialize();
        check(msa_batch_run(batch_->get(), nu_, tolerance, maxIterations, innerTol, maxOuter,
                            graphIters));
    }

    std::vector<double> Solution(int cls, int sample) const
    {
        std::vector<double> u(models_[cls]->TotalDOFs);
        check(msa_batch_get_solution(batch_->get(), cls, sample, u.data()));
        return u;
    }

    std::vector<int> Status(int cls) const
    {
        std::vector<int> st(samples_[cls]);
        check(msa_batch_get_stats(batch_->get(), cls, nullptr, nullptr, nullptr, st.data()));
        return st;
    }

  private:
    Context &ctx_;
    std::vector<const Model *> models_;
    std::vector<int> samples_;
    KLTable table_;
    std::vector<double> xi_;
    double meanE_, cov_, nu_;
    int precision_;
    std::vector<std::unique_ptr<Mesh>> meshes_;
    std::unique_ptr<Batch> batch_;
};

// MassAccelerationSinusoidalLoad (PreProcessor/MassAccelerationSinusoidalLoad.cs:
// 11-29): magnitude sin(2 pi / period * timeStep * i) for i < (int)(timeLength
// / timeStep), else 0.
class MassAccelerationSinusoidalLoad {
  public:
    MassAccelerationSinusoidalLoad(double period, double timeStep, double timeLength,
                                   double magnitude)
        : period_(period), timeStep_(timeStep), magnitude_(magnitude),
          timeSteps_((int)(timeLength / timeStep)) {}
    double operator[](int currentTimeStep) const
    {
        const double pi = 3.14159265358979323846;
        return currentTimeStep < timeSteps_
                   ? magnitude_ * std::sin(2 * pi / period_ * timeStep_ * currentTimeStep)
                   : 0.0;
    }
    DOFType DOF = DOFType::Z;

  private:
    double period_, timeStep_, magnitude_;
    int timeSteps_;
};

// NewmarkDynamicAnalyzer (Analyzers/NewmarkDynamicAnalyzer.cs:38-315) over a
// batch of stochastic-modulus samples: BuildMatrices() forms the effective
// stiffness (ProblemStructural.CalculateEffectiveMatrix), Solve() runs
// (int)(totalTime / timeStep) steps from rest with warm-started PCG.
class NewmarkDynamicAnalyzer {
  public:
    NewmarkDynamicAnalyzer(Context &ctx, const Model &model, std::vector<double> E, int samples,
                           double alpha, double delta, double timeStep, double totalTime,
                           double density, MassAccelerationSinusoidalLoad load,
                           int precision = MSA_MIXED, double rayleighAlpha = 0.0,
                           double rayleighBeta = 0.0, double nu = 0.3)
        : ctx_(ctx), model_(model), E_(std::move(E)), samples_(samples), alpha_(alpha),
          delta_(delta), timeStep_(timeStep), totalTime_(totalTime), density_(density),
          load_(load), precision_(precision), ra_(rayleighAlpha), rb_(rayleighBeta), nu_(nu) {}

    int Preconditioner = MSA_PRECOND_JACOBI;  // set before BuildMatrices()

    void BuildMatrices()
    {
        mesh_.reset(new Mesh(ctx_, model_));
        batch_.reset(new Batch(ctx_, {mesh_.get()}, {samples_}, precision_, 0));
        if (Preconditioner != MSA_PRECOND_JACOBI)
            check(msa_batch_set_preconditioner(batch_->get(), Preconditioner, 0, 0.0));
        check(msa_batch_set_modulus(batch_->get(), 0, E_.data()));
        check(msa_batch_set_rhs(batch_->get(), 0, model_.Forces.data()));
        check(msa_batch_set_dynamics(batch_->get(), density_, ra_, rb_, alpha_, delta_, timeStep_));
        check(msa_batch_set_probe(batch_->get(), 0, model_.TotalDOFs - 1));
        check(msa_batch_assemble(batch_->get(), nu_));
    }

    int Steps() const { return (int)(totalTime_ / timeStep_); }

    void Solve(double tolerance = 1e-8, int maxIterations = 20000, int graphIters = 16)
    {
        if (!batch_) BuildMatrices();
        std::vector<double> acc(Steps());
        for (int i = 0; i < Steps(); i++) acc[i] = load_[i];
        check(msa_batch_newmark(batch_->get(), Steps(), acc.data(), (int)load_.DOF, tolerance,
                                maxIterations, graphIters, 1));
    }

    // displacement of the last free DOF after every step
    std::vector<double> History(int sample) const
    {
        std::vector<double> h(Steps());
        check(msa_batch_get_history(batch_->get(), 0, sample, h.data()));
        return h;
    }

  private:
    Context &ctx_;
    const Model &model_;
    std::vector<double> E_;
    int samples_;
    double alpha_, delta_, timeStep_, totalTime_, density_;
    MassAccelerationSinusoidalLoad load_;
    int precision_;
    double ra_, rb_, nu_;
    std::unique_ptr<Mesh> mesh_;
    std::unique_ptr<Batch> batch_;
};

}  // namespace msolve
