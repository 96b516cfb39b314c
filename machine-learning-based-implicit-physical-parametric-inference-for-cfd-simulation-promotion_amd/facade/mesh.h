// Mesh facade: the reference's Mesh interface (IPPI-LBM/mesh.h:33-115) over
// the C ABI of libippi_lbm.so, so a driver written like IPPI-LBM/main.cpp
// runs the whole domain on one GPU. Public fields keep the reference's names
// and meaning; RankID must be 0 and CoreNum 1 (one GPU holds every column).
// The compile-time switches of IPPI-LBM/Parameter.h are runtime fields here
// (Collision, Precision, Boundary, UseLES).
#pragma once
#include <string>
#include <vector>

#include "../../include/ippi_lbm.h"

typedef double REAL;
struct Point2D {
    REAL x;
    REAL y;
};
// One point of the high-fidelity (XFlow) data (mesh.h:20-32): vx, vy, vz
// are the point's coordinates (lattice units after ReadXflowdata), ux, uy,
// uz its velocity, TotalPressure its density.
struct PointData {
    double VelocityModule;
    double StaticPress;
    double TotalPressure;
    double viscosity;
    double ux;
    double uy;
    double uz;
    double vx;
    double vy;
    double vz;
};

class Mesh {
public:
    Mesh();
    ~Mesh();
    // reference fields (mesh.h:37-52)
    int LatmeshSizeNx = 0;
    int LatmeshSizeNy = 0;
    Point2D center{0, 0};
    double radius = 0;
    double startTau = 0.6;
    Point2D startU{0, 0};
    int MyrankID = 0;
    int MyCoreNum = 1;
    int stepH = 0, stepL = 0;
    int startAvg = 0;
    double StartCs = 0.0;
    int errorflag = 0;
    int ParallelSize = 0;
    // runtime equivalents of Parameter.h:8-16
    int Collision = IPPI_COLL_BGK;   // SRTLBM / MRT / RRLBM
    int Precision = IPPI_FP64;       // REAL = double, or FP32 shifted populations
    int Boundary = IPPI_BC_INFLOW;   // IPPI_BC_FLATPLATE (the reference's own set), inflow + free slip, periodic
    int UseLES = 1;                  // SmagorinskyByHybrid
    int PlateX = 400;                // IPPI_BC_FLATPLATE: 'm' / 'd' split of the bottom wall (mesh.cpp:1413)
    uint64_t Seed = 0;
    double NoiseAmplitude = 0.0;
    // per-node relaxation times held by the facade (the reference's
    // LatMesh[i][j].tau), needed by FindAndWriteTau and ChangeTauFromXflow
    int TauField = 0;
    // boundary-row / column networks (NNmodel_Type01/03/05/07): weight file
    // per type, 101,249 little-endian FP64 values in nn.Linear order
    // (W0[128][15] b0, 6 x (W[128][128] b), W7[128] b7); empty = not used.
    // The reference loads TorchScript archives (AutoT0x.pt, mesh.cpp:2778).
    std::string NNmodelPath[8];
    // high-fidelity data files: XflowPrefix + (t + XflowStepOffset) + ".csv"
    std::string XflowPrefix = "Xflowdata_pingban_every_1_step/pb_";
    int XflowStepOffset = 10014;

    void InitialLatMesh(int Nx_P, int Ny, double tau, Point2D StartVel, double Rho, int RankID, int CoreNum,
                        int PartitionRemainderFlag);
    void InitialIBParticle(int NumNodes, double Radius, double ParticlePositionX, double ParticlePositionY,
                           Point2D ParticleVel);
    void OneLBMStep(int RankID, int CoreNum, REAL Cs, REAL Pr, int t);
    void JudgeStyle(Point2D CylinderCenter1, REAL Radius, int RankID, int CoreNum);
    void InitialFile();
    void WriteDatFluid(int t, int RankID, int NX, int CoreNum);
    void MergeDatFluid(int t, int CoreNum);
    void OutputCd(int RankID, int CoreNum, int time, double Radius);
    void ComputeAvgVelocity(int t);
    void WriteAvgu(int t, int RankID, int NX, int CoreNum);
    void MergeAvgu(int t, int CoreNum);
    void WriteDatParticle(int t);
    void ErrorCatch(int i, int j);
    // high-fidelity data and relaxation-time labels (mesh.cpp:2189-2766)
    std::vector<PointData> ReadXflowdata(int t);
    void ChangeTauFromXflow(int t);
    void FindAndWriteTau(int i, int j, int t);

    // direct access for callers that read the fields back
    int Fields(std::vector<double>& rho, std::vector<double>& ux, std::vector<double>& uy);
    ippi_handle* Handle() { return h_; }

private:
    void Check(int rc, const char* what);
    void EnsureHandle(REAL Cs);
    ippi_handle* h_ = nullptr;
    int nx_ = 0, ny_ = 0;
    double rho0_ = 1.0;
    Point2D u0_{0, 0};
    double cs_ = -1.0;
    bool fields_set_ = false;
    struct PendingBody {
        double cx, cy, d;
        int m;
        Point2D v;
    };
    std::vector<PendingBody> pending_;
    bool have_mean_ = false;
    int Means(std::vector<double>& rho, std::vector<double>& u, std::vector<double>& v);
    int last_force_rec_ = 0;
    std::vector<double> tau_;  // [y][x] when TauField
    void UploadTau();
    // index of the last data point within DX of each node (-1: none), the
    // reference's search order (mesh.cpp:2679-2691)
    std::vector<int> MatchXflow(const std::vector<PointData>& pts) const;
};
