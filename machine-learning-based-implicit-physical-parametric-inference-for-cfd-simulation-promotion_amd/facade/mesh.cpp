// Mesh facade over libippi_lbm.so (see mesh.h). Each method cites the
// reference method it stands in for (IPPI-LBM/mesh.cpp).
#include "mesh.h"

#include <sys/stat.h>

#include <cstdio>
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
    c.tau_mode = IPPI_TAU_CONST;
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
}

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
