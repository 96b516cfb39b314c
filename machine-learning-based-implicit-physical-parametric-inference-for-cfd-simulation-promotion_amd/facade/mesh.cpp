// Mesh facade over libippi_lbm.so (see mesh.h). Each method cites the
// reference method it stands in for (IPPI-LBM/mesh.cpp).
#include "mesh.h"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <map>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

Mesh::Mesh() {}
Mesh::~Mesh() {
    if (h_) ippi_destroy(h_);
}

void Mesh::Check(int rc, const char* what) {
    if (rc == IPPI_OK) return;
    if (rc == IPPI_ERR_DIVERGED) {  // ErrorCatch, mesh.cpp:2101-2111
        std::cerr << "Simulation Error!" << std::endl;
        errorflag = 1;
        std::exit(1);
    }
    std::cerr << what << " failed (" << rc << "): " << (h_ ? ippi_last_error(h_) : "") << std::endl;
    std::exit(2);
}

// InitialLatMesh, mesh.cpp:60-171: the whole domain is one slab, so
// Nx_P must be the global Nx (RankID 0, CoreNum 1, no remainder).
void Mesh::InitialLatMesh(int Nx_P, int Ny, double tau, Point2D StartVel, double Rho, int RankID, int CoreNum,
                          int PartitionRemainderFlag) {
    if (RankID != 0 || CoreNum != 1 || PartitionRemainderFlag != 0) {
        std::cerr << "Mesh facade: one GPU holds the whole domain (RankID 0, CoreNum 1)" << std::endl;
        std::exit(2);
    }
    nx_ = Nx_P;
    ny_ = Ny;
    if (LatmeshSizeNx == 0) LatmeshSizeNx = Nx_P;  // main.cpp:64-65 sets them; one rank holds all columns
    if (LatmeshSizeNy == 0) LatmeshSizeNy = Ny;
    startTau = tau;
    rho0_ = Rho;
    u0_ = StartVel;
    ParallelSize = Nx_P;
}

void Mesh::EnsureHandle(REAL Cs) {
    if (h_ && cs_ == Cs) return;
    if (h_) {
        std::cerr << "Mesh facade: Cs must stay constant over the run" << std::endl;
        std::exit(2);
    }
    ippi_config c{};
    c.nx = nx_;
    c.ny = ny_;
    c.collision = Collision;
    c.precision = Precision;
    c.bc = Boundary;
    c.les = UseLES;
    c.tau_mode = TauField ? IPPI_TAU_FIELD : IPPI_TAU_CONST;
    c.tau0 = startTau;
    c.cs = Cs;
    c.u_in_x = startU.x;
    c.u_in_y = startU.y;
    c.plate_x = PlateX;
    Check(ippi_create(&c, &h_), "ippi_create");
    cs_ = Cs;
    Check(ippi_init_fields(h_, Seed, rho0_, u0_.x, u0_.y, NoiseAmplitude), "ippi_init_fields");
    for (const auto& b : pending_) {
        int id = 0;
        Check(ippi_add_body(h_, b.cx, b.cy, b.d, b.m, b.v.x, b.v.y, &id), "ippi_add_body");
    }
    pending_.clear();
    if (TauField) {
        tau_.assign((size_t)nx_ * ny_, startTau);
        UploadTau();
    }
    // boundary networks: loaded once, evaluated by every OneLBMStep (the
    // reference reloads its model file on every call, mesh.cpp:2780)
    for (int type = 1; type <= 7; type += 2) {
        if (NNmodelPath[type].empty()) continue;
        std::vector<double> w(101249);
        std::FILE* fp = std::fopen(NNmodelPath[type].c_str(), "rb");
        const size_t got = fp ? std::fread(w.data(), sizeof(double), w.size(), fp) : 0;
        if (fp) std::fclose(fp);
        if (got != w.size()) {
            std::cerr << "Mesh facade: cannot read 101249 values from " << NNmodelPath[type] << std::endl;
            std::exit(2);
        }
        Check(ippi_set_boundary_mlp(h_, type, w.data(), (int)w.size()), "ippi_set_boundary_mlp");
    }
}

void Mesh::UploadTau() { Check(ippi_set_tau_field(h_, tau_.data()), "ippi_set_tau_field"); }

// InitialIBParticle, mesh.cpp:173-191 (markers on a circle of radius Radius)
void Mesh::InitialIBParticle(int NumNodes, double Radius, double X, double Y, Point2D Vel) {
    PendingBody b{X, Y, 2.0 * Radius, NumNodes, Vel};
    if (h_) {
        int id = 0;
        Check(ippi_add_body(h_, b.cx, b.cy, b.d, b.m, b.v.x, b.v.y, &id), "ippi_add_body");
    } else {
        pending_.push_back(b);
    }
}

// JudgeStyle, mesh.cpp:1391-1454: node classes follow from the boundary
// kind chosen in Boundary; nothing is stored per node.
void Mesh::JudgeStyle(Point2D, REAL, int, int) {}

// OneLBMStep, mesh.cpp:357-513 (IB forcing included when bodies exist)
void Mesh::OneLBMStep(int RankID, int CoreNum, REAL Cs, REAL Pr, int t) {
    (void)RankID; (void)CoreNum; (void)Pr; (void)t;
    EnsureHandle(Cs);
    Check(ippi_step(h_, 1), "ippi_step");
}

void Mesh::ErrorCatch(int, int) {
    if (h_) Check(ippi_synchronize(h_), "ippi_synchronize");
}

int Mesh::Fields(std::vector<double>& rho, std::vector<double>& ux, std::vector<double>& uy) {
    const size_t N = (size_t)nx_ * ny_;
    rho.resize(N);
    ux.resize(N);
    uy.resize(N);
    return ippi_get_fields(h_, rho.data(), ux.data(), uy.data());
}

// InitialFile, mesh.cpp:1171-1179
void Mesh::InitialFile() {
    mkdir("fluid", 0755);
    mkdir("particle", 0755);
    mkdir("cp", 0755);
}

// ComputeAvgVelocity, mesh.cpp:1954-1974: running means from t = startAvg,
// avg = (avg (t - startAvg) + x) / (t - startAvg + 1), kept on the device.
void Mesh::ComputeAvgVelocity(int t) {
    if (t < startAvg || !h_) return;
    Check(ippi_accumulate_mean(h_, (int64_t)(t - startAvg)), "ippi_accumulate_mean");
    have_mean_ = true;
}

// Means for the writers (zeros before the first ComputeAvgVelocity, as the
// reference's avg fields start at zero).
int Mesh::Means(std::vector<double>& rho, std::vector<double>& u, std::vector<double>& v) {
    const size_t N = (size_t)nx_ * ny_;
    rho.assign(N, 0.0);
    u.assign(N, 0.0);
    v.assign(N, 0.0);
    if (!have_mean_) return IPPI_OK;
    return ippi_get_mean(h_, rho.data(), u.data(), v.data());
}

// WriteDatFluid, mesh.cpp:1180-1243: Tecplot ASCII, point packing, x-major
// (I = Ny fastest), variables X Y Density U V avgu avgv tau avgRho; tau is
// the relaxation time of each node's next collision.
void Mesh::WriteDatFluid(int t, int RankID, int NX, int CoreNum) {
    (void)CoreNum;
    std::vector<double> r, u, v, ar, au, av, tau((size_t)nx_ * ny_, startTau);
    Check(Fields(r, u, v), "ippi_get_fields");
    Check(Means(ar, au, av), "ippi_get_mean");
    if (ippi_get_tau(h_, tau.data()) != IPPI_OK) tau.assign(tau.size(), startTau);  // before the first step
    std::stringstream fn;
    fn << "fluid/fluid_t" << t << "_" << RankID << ".dat";
    std::ofstream out(fn.str());
    out << "TITLE = \"Simulation of Fluid State at Time " << t << "\"\n";
    out << "VARIABLES = \"X\" \"Y\" \"Density\" \"U\" \"V\" \"avgu\" \"avgv\" \"tau\" \"avgRho\" \n";
    out << "ZONE T=\"Time " << t << "\", ";
    out << "I=" << ny_ << ", J=" << NX << ", K=1, ";
    out << "DATAPACKING=POINT, ZONETYPE=Ordered\n";
    for (int X = 0; X < nx_; ++X)
        for (int Y = 0; Y < ny_; ++Y) {
            const size_t n = (size_t)Y * nx_ + X;
            out << X << " " << Y << " " << r[n] << " " << u[n] << " " << v[n] << " " << au[n] << " " << av[n] << " "
                << tau[n] << " " << ar[n] << "\n";
        }
}

// WriteAvgu, mesh.cpp:2009-2067: rows 1 .. Ny-2 only (I = Ny-2), variables
// X Y AvgdDenstiy U V avgy avgv Style; the density column holds rho (as the
// reference writes it); Style 0 for fluid nodes, 3 for inlet / outlet.
void Mesh::WriteAvgu(int t, int RankID, int NX, int CoreNum) {
    (void)CoreNum;
    std::vector<double> r, u, v, ar, au, av;
    Check(Fields(r, u, v), "ippi_get_fields");
    Check(Means(ar, au, av), "ippi_get_mean");
    std::stringstream fn;
    fn << "fluid/Avgu_" << t << "_" << RankID << ".dat";
    std::ofstream out(fn.str());
    out << "TITLE = \"Simulation of Fluid State at Time " << t << "\"\n";
    out << "VARIABLES = \"X\" \"Y\" \"AvgdDenstiy\" \"U\" \"V\" \"avgy\" \"avgv\" \"Style\" \n";
    out << "ZONE T=\"Time " << t << "\", ";
    out << "I=" << ny_ - 2 << ", J=" << NX << ", K=1, ";
    out << "DATAPACKING=POINT, ZONETYPE=Ordered\n";
    const bool nee_cols = Boundary != IPPI_BC_PERIODIC_BB;
    for (int X = 0; X < nx_; ++X)
        for (int Y = 1; Y < ny_ - 1; ++Y) {
            const size_t n = (size_t)Y * nx_ + X;
            const int style = nee_cols && (X == 0 || X == nx_ - 1) ? 3 : 0;
            out << X << " " << Y << " " << r[n] << " " << u[n] << " " << v[n] << " " << au[n] << " " << av[n] << " "
                << style << "\n";
        }
}

// MergeAvgu, mesh.cpp:2069-2099: one part file with one GPU.
void Mesh::MergeAvgu(int t, int CoreNum) {
    (void)CoreNum;
    std::stringstream src, dst;
    src << "fluid/Avgu_" << t << "_0.dat";
    dst << "fluid/fluid_avg_" << t << ".dat";
    std::rename(src.str().c_str(), dst.str().c_str());
}

// WriteDatParticle, mesh.cpp:1283-1319: Tecplot FE line segments through the
// markers of every body, closed by an element from node 1 to the last node.
void Mesh::WriteDatParticle(int t) {
    int n = 0;
    Check(ippi_get_markers(h_, nullptr, 1 << 30, &n), "ippi_get_markers");  // count
    std::vector<double> mk(6 * (size_t)(n > 0 ? n : 1));
    Check(ippi_get_markers(h_, mk.data(), n, &n), "ippi_get_markers");
    std::stringstream fn;
    fn << "particle/particle_t" << t << ".dat";
    std::ofstream out(fn.str());
    out << "TITLE = \"Particle State at Time " << t << "\"\n";
    out << "VARIABLES = \"X\" \"Y\" \"Z\"\n";
    out << "ZONE T=\"Particle\", N=" << n << ", E=" << n;
    out << ", DATAPACKING=POINT, ZONETYPE=FELINESEG\n";
    out << "Nodes=" << n << ", Elements=" << n << "\n";
    for (int m = 0; m < n; ++m) out << mk[6 * m] << " " << mk[6 * m + 1] << " 0\n";
    for (int m = 0; m + 1 < n; ++m) out << m + 1 << " " << m + 2 << "\n";
    out << 1 << " " << n << "\n";
}

// MergeDatFluid, mesh.cpp:1320-1353: with one GPU there is one part file.
void Mesh::MergeDatFluid(int t, int CoreNum) {
    (void)CoreNum;
    std::stringstream src, dst;
    src << "fluid/fluid_t" << t << "_0.dat";
    dst << "fluid/fluid_Merged_" << t << ".dat";
    std::rename(src.str().c_str(), dst.str().c_str());
}

// OutputCd, mesh.cpp:1461-1526: drag/lift coefficients of body 0 from the
// last IB force record, C = F / (0.5 U^2 2R), appended to particle/CD_CL.dat.
void Mesh::OutputCd(int RankID, int CoreNum, int time, double Radius) {
    (void)CoreNum;
    int nb = 0, n = 0;
    Check(ippi_num_bodies(h_, &nb), "ippi_num_bodies");
    if (nb == 0) return;
    std::vector<double> rec(2 * nb);
    Check(ippi_get_body_forces(h_, rec.data(), 1, &n), "ippi_get_body_forces");
    if (n == 0) return;
    double cd = rec[0] / (0.5 * startU.x * startU.x * 2 * Radius);
    double cl = rec[1] / (0.5 * startU.x * startU.x * 2 * Radius);
    if (RankID == 0) {
        std::cerr << "cd:" << cd << " " << "cl:" << cl << std::endl;
        std::ofstream f("particle/CD_CL.dat", std::fstream::app);
        f << time << " " << cd << " " << cl << std::endl;
    }
}

// ---------------------------------------------------------------------------
// High-fidelity data (XFlow CSV) and relaxation-time labels.

// ReadXflowdata, mesh.cpp:2414-2545: one CSV per step, header line skipped;
// each line is read as the reference's stream extraction reads it (13
// comma-terminated numbers: id, x, y, z, VelocityModule, StaticPressure, -,
// TotalPressure, -, viscosity, u, v, w; a line that does not parse is
// skipped). Coordinates are rounded to 1e-3 m, then converted to lattice
// units (x + 0.4 m, dx = 1 mm), velocities by the lattice sound speed
// (340.112 m/s / sqrt 3), density from the static pressure; points with the
// same (x, y) are merged by halving after each addition; points outside the
// columns this rank holds are dropped. Merged points are kept in the order
// of their first occurrence.
std::vector<PointData> Mesh::ReadXflowdata(int t) {
    const std::string name = XflowPrefix + std::to_string(t + XflowStepOffset) + ".csv";
    std::ifstream file(name);
    std::vector<PointData> data;
    if (!file.is_open()) {
        std::cerr << "ReadXflowdata: cannot open " << name << std::endl;
        return data;
    }
    std::string line;
    std::getline(file, line);  // header
    while (std::getline(file, line)) {
        std::stringstream ss(line);
        PointData p{};
        char comma;
        double unused;
        if (!(ss >> unused >> comma >> p.vx >> comma >> p.vy >> comma >> p.vz >> comma >> p.VelocityModule >> comma >>
              p.StaticPress >> comma >> unused >> comma >> p.TotalPressure >> comma >> unused >> comma >>
              p.viscosity >> comma >> p.ux >> comma >> p.uy >> comma >> p.uz >> comma))
            continue;
        p.vx = std::round(p.vx * 1000.0) / 1000.0;  // RoundToThreeDecimalPlaces, mesh.cpp:41-46
        p.vy = std::round(p.vy * 1000.0) / 1000.0;
        data.push_back(p);
    }
    const double DX = 0.001, dv = 340.112;
    for (auto& p : data) {
        p.vx = p.vx + 0.4;
        p.vx = p.vx / DX;
        p.vy = p.vy / DX;
        p.vz = p.vz / DX;
        p.VelocityModule = p.VelocityModule / dv / std::sqrt(3.0);
        p.ux = p.ux / dv / std::sqrt(3.0);
        p.uy = p.uy / dv / std::sqrt(3.0);
        p.uz = p.uz / dv / std::sqrt(3.0);
        p.TotalPressure = (p.StaticPress + 101325) / 101325;
    }
    std::vector<PointData> merged;
    std::map<std::pair<double, double>, size_t> seen;
    for (const auto& p : data) {
        auto it = seen.find({p.vx, p.vy});
        if (it == seen.end()) {
            seen[{p.vx, p.vy}] = merged.size();
            merged.push_back(p);
            continue;
        }
        PointData& c = merged[it->second];
        c.VelocityModule = (c.VelocityModule + p.VelocityModule) / 2.0;
        c.StaticPress = (c.StaticPress + p.StaticPress) / 2.0;
        c.TotalPressure = (c.TotalPressure + p.TotalPressure) / 2.0;
        c.ux = (c.ux + p.ux) / 2.0;
        c.uy = (c.uy + p.uy) / 2.0;
        c.uz = (c.uz + p.uz) / 2.0;
        c.vz = (c.vz + p.vz) / 2.0;
        c.viscosity = (c.viscosity + p.viscosity) / 2.0;
    }
    // columns of this rank (one rank: the whole domain and the two ghosts)
    const double lo = -2.0 - 0.5, hi = (double)LatmeshSizeNx + 0.5;
    std::vector<PointData> out;
    for (const auto& p : merged)
        if (!(p.vx < lo || p.vx > hi)) out.push_back(p);
    return out;
}

std::vector<int> Mesh::MatchXflow(const std::vector<PointData>& pts) const {
    const double DX = 0.001;
    std::vector<int> m((size_t)nx_ * ny_, -1);
    for (size_t q = 0; q < pts.size(); q++) {
        const double X = std::round(pts[q].vx), Y = std::round(pts[q].vy);
        if (!(std::abs(pts[q].vx - X) < DX && std::abs(pts[q].vy - Y) < DX)) continue;
        if (X < 0 || X >= nx_ || Y < 0 || Y >= ny_) continue;
        m[(size_t)Y * nx_ + (size_t)X] = (int)q;  // the last point in the search order wins
    }
    return m;
}

// ChangeTauFromXflow, mesh.cpp:2730-2766: tau = 0.5 + 3 nu from the data at
// every node, in the reference's node order; a node without data keeps the
// viscosity and misfits of the node before it (the first ones: startTau and
// zero misfits). The mean squared velocity and density misfits over the
// nodes go to cp/InletMSE.dat.
void Mesh::ChangeTauFromXflow(int t) {
    if (!h_ || !TauField) {
        std::cerr << "ChangeTauFromXflow: needs TauField = 1 and an initialised mesh" << std::endl;
        std::exit(2);
    }
    const std::vector<PointData> pts = ReadXflowdata(t);
    const std::vector<int> m = MatchXflow(pts);
    std::vector<double> rho, ux, uy;
    Check(Fields(rho, ux, uy), "ippi_get_fields");
    double nu = (startTau - 0.5) / 3.0, ev = 0.0, er = 0.0, sv = 0.0, sr = 0.0;
    for (int x = 0; x < nx_; x++)
        for (int y = 0; y < ny_; y++) {
            const size_t n = (size_t)y * nx_ + x;
            if (m[n] >= 0) {
                const PointData& p = pts[(size_t)m[n]];
                nu = p.viscosity;
                ev = (ux[n] - p.ux) * (ux[n] - p.ux) + (uy[n] - p.uy) * (uy[n] - p.uy);
                er = (rho[n] - p.TotalPressure) * (rho[n] - p.TotalPressure);
            }
            sv += ev;
            sr += er;
            tau_[n] = 0.5 + 3.0 * nu;
        }
    UploadTau();
    std::ofstream out("cp/InletMSE.dat", std::ios::app);
    out << t << " " << sv / ((double)nx_ * ny_) << " " << sr / ((double)nx_ * ny_) << std::endl;
}

// FindAndWriteTau, mesh.cpp:2647-2728: the relaxation-time shift of the nine
// nodes around (i, j) that best reproduces the data of step t + 1 at (i, j):
// the candidate search runs on the GPU (ippi_tau_search) over the inputs of
// the last step's collision (ippi_get_stencil); tau += the best shift on the
// nine nodes, and a line "t best tau_0 .. tau_8" goes to cp/BestTau.dat. i is
// the reference's slab column (ghost column 0), so the node is x = i - 1.
void Mesh::FindAndWriteTau(int i, int j, int t) {
    if (!h_ || !TauField) {
        std::cerr << "FindAndWriteTau: needs TauField = 1 and an initialised mesh" << std::endl;
        std::exit(2);
    }
    const int x = i - 1, y = j;
    const std::vector<PointData> pts = ReadXflowdata(t + 1);
    const std::vector<int> m = MatchXflow(pts);
    if (x < 0 || x >= nx_ || y < 0 || y >= ny_ || m[(size_t)y * nx_ + x] < 0) {
        std::cerr << "FindAndWriteTau: no data at node (" << x << ", " << y << ")" << std::endl;
        return;
    }
    const PointData& p = pts[(size_t)m[(size_t)y * nx_ + x]];
    const int xy[2] = {x, y};
    double fn[9], fq[9], tk[9], target[3] = {p.TotalPressure, p.ux, p.uy}, best = 0.0, err = 0.0;
    Check(ippi_get_stencil(h_, 1, xy, fn, fq, tk), "ippi_get_stencil");
    Check(ippi_tau_search(0, 1, fn, fq, tk, target, -10000, 10000, 1e-5, &best, &err), "ippi_tau_search");
    static const int EXk[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1}, EYk[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    for (int k = 0; k < 9; k++) tau_[(size_t)(y + EYk[k]) * nx_ + (x + EXk[k])] += best;
    UploadTau();
    std::cerr << "Timestep:" << t << " " << "BestTau:" << best << std::endl;
    std::ofstream out("cp/BestTau.dat", std::ios::app);
    out << t << " " << best << " ";
    for (int k = 0; k < 9; k++) out << tau_[(size_t)(y + EYk[k]) * nx_ + (x + EXk[k])] << " ";
    out << std::endl;
}
