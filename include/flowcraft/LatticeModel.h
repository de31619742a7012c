// Minimal stand-alone definition of FlowCraft's LatticeModel enum
// (FlowCraft/LatticeModel.h:12) for builds of LbmSolver.h outside FlowCraft; inside
// FlowCraft the application's own header is found first on the include path.
#pragma once

enum class LatticeModel { D3Q15, D3Q19 };
