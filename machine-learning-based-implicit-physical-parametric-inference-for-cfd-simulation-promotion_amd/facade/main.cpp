// Driver written like IPPI-LBM/main.cpp:3-229, on the Mesh facade: one GPU
// holds the whole domain (RankID 0, CoreNum 1). Workload: a cylinder of
// diameter 80 lu (251 markers) in a 2048 x 1024 channel, MRT + hybrid
// Smagorinsky LES, Re = 1e4, inflow 0.08; prints cd/cl every SaveStep.
//   usage: ippi_main [NumStep] [SaveStep]
#include <cstdlib>
#include <iostream>

#include "mesh.h"

int main(int argc, char* argv[]) {
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
    const int NumStep = argc > 1 ? std::atoi(argv[1]) : 2000;
    const int SaveStep = argc > 2 ? std::atoi(argv[2]) : 500;

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
