// Configuration enums, same enumerators and order as the reference's
// include/pmConfig.h:3-5 (their values are the PMGPU_* constants of pmgpu.h).
#pragma once

enum class InterpolationScheme { NGP, CIC, TSC };
enum class FiniteDiffScheme { TWO_POINT, FOUR_POINT };
enum class GreensFunction { DISCRETE_LAPLACIAN, S1_OPTIMAL, S2_OPTIMAL, POOR_MAN };
