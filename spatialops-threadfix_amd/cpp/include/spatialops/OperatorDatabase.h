// OperatorDatabase.h -- OperatorDatabase and build_stencils of the C++ front end.
#pragma once
#include <spatialops/NeboStencilBuilder.h>
