// Drivers written like IPPI-LBM/main.cpp:3-229, on the Mesh facade: one GPU
// holds the whole domain (RankID 0, CoreNum 1).
//   ippi_main [NumStep] [SaveStep]
//       cylinder of diameter 80 lu (251 markers) in a 2048 x 1024 channel,
//       MRT + hybrid Smagorinsky LES, Re = 1e4, inflow 0.08, FP32 shifted
//       populations; appends cd/cl to particle/CD_CL.dat every SaveStep.
//   ippi_main check
//       every writer on a small case (see run_check).
//   ippi_main xflow <prefix>
//       high-fidelity data and relaxation-time labels (see run_xflow).
//   ippi_main plate [NumStep] [SaveStep]
//       the reference's own run (main.cpp:13-48): flat plate on 2401 x 501,
//       BGK + hybrid Smagorinsky (Cs = 0.173), U = 0.1, Re = 5e5,
//       tau = 0.5 + 6 Ny U / Re, FP64, the reference's boundary set
//       (free slip ahead of x = 400, no slip beyond); Tecplot fields every
//       SaveStep (WriteDatFluid + MergeDatFluid), running means from
//       startAvg = 30000 (ComputeAvgVelocity).
//   ippi_main plate NumStep SaveStep NNdir
//       the same with the boundary-row / column networks whose weight files
//       NNdir/AutoT0{1,3,5,7}.bin exist (101,249 FP64 values each).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "mesh.h"

static int run_cylinder(int NumStep, int SaveStep) {
    int RankID = 0, CoreNum = 1;
    REAL Cs = 0.1;
    REAL Pr = 1;
    int Nx = 2048, Ny = 1024;
    const int ParticleNumNodes = 251;
    const REAL Radius = 40;
    Point2D CylinderCenter{512, 512};
    Point2D ParticleVel{0.0, 0.0};
    double Re = 1e4;
    Point2D StartVel{0.08, 0.0};
    const REAL Rho = 1.0;
    const REAL tau = 0.5 + 3.0 * StartVel.x * (2 * Radius) / Re;

    Mesh mesh;
    mesh.startAvg = NumStep;
    mesh.center = CylinderCenter;
    mesh.radius = Radius;
    mesh.startTau = tau;
    mesh.startU = StartVel;
    mesh.MyrankID = RankID;
    mesh.MyCoreNum = CoreNum;
    mesh.StartCs = Cs;
    mesh.LatmeshSizeNx = Nx;
    mesh.LatmeshSizeNy = Ny;
    mesh.Collision = IPPI_COLL_MRT;
    mesh.Precision = IPPI_FP32_SHIFTED;
    mesh.Boundary = IPPI_BC_INFLOW;
    mesh.Seed = 20240601;
    mesh.NoiseAmplitude = 1e-3;
    mesh.InitialFile();
    mesh.InitialLatMesh(Nx, Ny, tau, StartVel, Rho, RankID, CoreNum, 0);
    mesh.JudgeStyle(CylinderCenter, Radius, RankID, CoreNum);
    mesh.InitialIBParticle(ParticleNumNodes, Radius, CylinderCenter.x, CylinderCenter.y, ParticleVel);
    std::cout << "Initial Completed" << std::endl;
    std::cout << "Re = " << Re << " Tau = " << tau << std::endl;
    for (int t = 1; t <= NumStep; ++t) {
        mesh.OneLBMStep(RankID, CoreNum, Cs, Pr, t);
        if (t % SaveStep == 0) {
            mesh.OutputCd(RankID, CoreNum, t, Radius);
            std::cout << "Compute Step: " << t << std::endl;
        }
    }
    mesh.ErrorCatch(0, 0);
    return 0;
}

static int run_plate(int NumStep, int SaveStep, const char* NNdir) {
    int RankID = 0, CoreNum = 1;
    REAL Cs = 0.173;
    REAL Pr = 1;
    int Nx = 2401;
    int Ny = 501;
    int StepH = 0, StepL = 0;
    const REAL Radius = 20;
    Point2D CylinderCenter{400, 620};
    double Re = 5e5;
    Point2D StartVel{0.1, 0.0};
    const REAL Rho = 1.0;
    const REAL tau = 0.5 + 6.0 * ((Ny * StartVel.x) / Re);

    Mesh mesh;
    mesh.startAvg = 30000;
    mesh.center = CylinderCenter;
    mesh.radius = Radius;
    mesh.startTau = tau;
    mesh.startU = StartVel;
    mesh.MyrankID = RankID;
    mesh.MyCoreNum = CoreNum;
    mesh.stepH = StepH;
    mesh.stepL = StepL;
    mesh.StartCs = Cs;
    mesh.LatmeshSizeNx = Nx;
    mesh.LatmeshSizeNy = Ny;
    mesh.Collision = IPPI_COLL_BGK;
    mesh.Precision = IPPI_FP64;
    mesh.Boundary = IPPI_BC_FLATPLATE;
    mesh.PlateX = 400;
    // optional boundary networks: <NNdir>/AutoT0{1,3,5,7}.bin where present
    // (the reference's AutoT0x.pt models, mesh.cpp:2778, as raw FP64 weights)
    if (NNdir) {
        for (int type = 1; type <= 7; type += 2) {
            const std::string fn = std::string(NNdir) + "/AutoT0" + std::to_string(type) + ".bin";
            if (std::FILE* fp = std::fopen(fn.c_str(), "rb")) {
                std::fclose(fp);
                mesh.NNmodelPath[type] = fn;
            }
        }
    }
    mesh.InitialFile();
    mesh.InitialLatMesh(Nx, Ny, tau, StartVel, Rho, RankID, CoreNum, 0);
    mesh.JudgeStyle(CylinderCenter, Radius, RankID, CoreNum);
    mesh.ParallelSize = Nx;
    std::cout << "Initial Completed" << std::endl;
    std::cout << "U = " << StartVel.x << " Re = " << Re << " Tau = " << tau << std::endl;
    for (int t = 1; t <= NumStep; ++t) {
        mesh.OneLBMStep(RankID, CoreNum, Cs, Pr, t);
        if (t % SaveStep == 0) {
            mesh.WriteDatFluid(t, RankID, Nx, CoreNum);
            mesh.MergeDatFluid(t, CoreNum);
            std::cout << "Compute Step: " << t << std::endl;
        }
        mesh.ComputeAvgVelocity(t);
    }
    mesh.ErrorCatch(0, 0);
    return 0;
}

// Every writer of the facade on a small case (128 x 64, cylinder D = 16,
// BGK, periodic x / bounce-back walls, FP64): running means from step 20,
// then at step 50 WriteDatFluid + MergeDatFluid, WriteAvgu + MergeAvgu,
// WriteDatParticle and OutputCd.
static int run_check() {
    int RankID = 0, CoreNum = 1;
    const int Nx = 128, Ny = 64, NumStep = 50;
    Point2D StartVel{0.05, 0.0};
    Mesh mesh;
    mesh.startAvg = 20;
    mesh.startTau = 0.6;
    mesh.startU = StartVel;
    mesh.Collision = IPPI_COLL_BGK;
    mesh.Precision = IPPI_FP64;
    mesh.Boundary = IPPI_BC_PERIODIC_BB;
    mesh.UseLES = 0;
    mesh.Seed = 7;
    mesh.NoiseAmplitude = 1e-3;
    mesh.InitialFile();
    mesh.InitialLatMesh(Nx, Ny, 0.6, StartVel, 1.0, RankID, CoreNum, 0);
    mesh.InitialIBParticle(51, 8.0, 40.0, 32.0, Point2D{0.0, 0.0});
    for (int t = 1; t <= NumStep; ++t) {
        mesh.OneLBMStep(RankID, CoreNum, 0.0, 1.0, t);
        mesh.ComputeAvgVelocity(t);
    }
    mesh.WriteDatFluid(NumStep, RankID, Nx, CoreNum);
    mesh.MergeDatFluid(NumStep, CoreNum);
    mesh.WriteAvgu(NumStep, RankID, Nx, CoreNum);
    mesh.MergeAvgu(NumStep, CoreNum);
    mesh.WriteDatParticle(NumStep);
    mesh.OutputCd(RankID, CoreNum, NumStep, 8.0);
    mesh.ErrorCatch(0, 0);
    std::cout << "Check Completed" << std::endl;
    return 0;
}

// High-fidelity data and relaxation-time labels on a small case (64 x 40,
// FP64 BGK + LES, inflow, per-node tau): data files prefix + (t + 10014) +
// ".csv". Writes the parsed points of step 10 and the inputs of
// FindAndWriteTau (its result goes to cp/BestTau.dat) at
// the reference's slab column i = 21 (node x = 20), row 20, step 15.
static int run_xflow(const char* prefix) {
    int RankID = 0, CoreNum = 1;
    const int Nx = 64, Ny = 40;
    Mesh mesh;
    mesh.LatmeshSizeNx = Nx;
    mesh.LatmeshSizeNy = Ny;
    mesh.Collision = IPPI_COLL_BGK;
    mesh.Precision = IPPI_FP64;
    mesh.Boundary = IPPI_BC_INFLOW;
    mesh.UseLES = 1;
    mesh.TauField = 1;
    mesh.Seed = 5;
    mesh.NoiseAmplitude = 1e-3;
    mesh.XflowPrefix = prefix;
    mesh.InitialFile();
    mesh.InitialLatMesh(Nx, Ny, 0.58, Point2D{0.05, 0.0}, 1.0, RankID, CoreNum, 0);
    for (int t = 1; t <= 10; t++) mesh.OneLBMStep(RankID, CoreNum, 0.12, 1.0, t);
    std::FILE* f = std::fopen("xflow_points.txt", "w");
    for (const auto& p : mesh.ReadXflowdata(10))
        std::fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", p.vx, p.vy, p.ux, p.uy, p.TotalPressure,
                     p.viscosity, p.VelocityModule);
    std::fclose(f);
    mesh.ChangeTauFromXflow(10);
    for (int t = 11; t <= 15; t++) mesh.OneLBMStep(RankID, CoreNum, 0.12, 1.0, t);
    const int xy[2] = {20, 20};
    double fn[9], fq[9], tk[9];
    if (ippi_get_stencil(mesh.Handle(), 1, xy, fn, fq, tk) != IPPI_OK) return 1;
    f = std::fopen("xflow_stencil.txt", "w");
    for (int k = 0; k < 9; k++) std::fprintf(f, "%.17g %.17g %.17g\n", fn[k], fq[k], tk[k]);
    std::fclose(f);
    mesh.FindAndWriteTau(21, 20, 15);
    std::cout << "Xflow Completed" << std::endl;
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc > 1 && std::strcmp(argv[1], "check") == 0) return run_check();
    if (argc > 2 && std::strcmp(argv[1], "xflow") == 0) return run_xflow(argv[2]);
    if (argc > 1 && std::strcmp(argv[1], "plate") == 0)
        return run_plate(argc > 2 ? std::atoi(argv[2]) : 200000, argc > 3 ? std::atoi(argv[3]) : 20000,
                         argc > 4 ? argv[4] : nullptr);
    return run_cylinder(argc > 1 ? std::atoi(argv[1]) : 2000, argc > 2 ? std::atoi(argv[2]) : 500);
}
