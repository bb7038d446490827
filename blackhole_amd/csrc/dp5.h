// dp5.h - Dormand-Prince 5(4) constants shared by the tracer and the adjoint.
#pragma once

namespace bhx {
namespace dp5 {

// Dormand-Prince 5(4) tableau, error weights and Hairer's dense-output
// weights (Hairer, Norsett & Wanner, Solving ODEs I, II.4).
constexpr double A21 = 1.0 / 5.0;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                 A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                 A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
constexpr double B1 = 35.0 / 384.0, B3 = 500.0 / 1113.0, B4 = 125.0 / 192.0,
                 B5 = -2187.0 / 6784.0, B6 = 11.0 / 84.0;
constexpr double E1 = B1 - 5179.0 / 57600.0, E3 = B3 - 7571.0 / 16695.0,
                 E4 = B4 - 393.0 / 640.0, E5 = B5 - (-92097.0 / 339200.0),
                 E6 = B6 - 187.0 / 2100.0, E7 = 0.0 - 1.0 / 40.0;
constexpr double D1 = -12715105075.0 / 11282082432.0, D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0, D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0, D7 = 69997945.0 / 29380423.0;
// PI controller exponents alpha = 0.7/5, beta = 0.4/5, safety 0.9 (geodesic.h:170).
constexpr double kAlpha = 0.7 / 5.0, kBeta = 0.4 / 5.0;
constexpr double kLog1em4 = -9.210340371976182;  // log(1e-4): err_prev floor

}  // namespace dp5
}  // namespace bhx
