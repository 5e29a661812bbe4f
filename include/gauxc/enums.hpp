// Enumerations of the GauXC-compatible C++ surface (include/gauxc/enums.hpp
// in GauXC).  Values follow the order of the GauXC C enums
// (include/gauxc/c/enums.h:24-84) so C and C++ callers agree.
#pragma once

namespace GauXC {

enum class RadialQuad { Becke, MuraKnowles, MurrayHandyLaming, TreutlerAhlrichs };

enum class AtomicGridSizeDefault { FineGrid, UltraFineGrid, SuperFineGrid, GM3, GM5 };

enum class XCWeightAlg { NOTPARTITIONED, Becke, SSF, LKO };

enum class ExecutionSpace { Host, Device };

enum class PruningScheme { Unpruned, Robust, Treutler };

}  // namespace GauXC
