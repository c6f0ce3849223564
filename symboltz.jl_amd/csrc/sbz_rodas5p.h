// Rodas5P tableau (Steinebach 2023, BIT 63:27; cited by the reference at paper/aa57450-25.tex:581)
// in the W-transformed form used by OrdinaryDiffEqRosenbrock, which the reference selects at
// src/solve.jl:320,339:
//   (I/(γh) - J) k_i = f(t + c_i h, u + Σ_{j<i} a_ij k_j) + Σ_{j<i} (C_ij/h) k_j + d_i h ∂f/∂t
//   u_{n+1} = u + Σ_{j<8} a_8j k_j + k_8,  error estimate k_8.
// Stages 7 and 8 reuse row 6 of `a` (a_7j = a_6j, a_76 = 1; a_8j = a_6j, a_86 = a_87 = 1).
// The coefficients are verified by the oracle tests through the local convergence order (6 for
// the solution, 5 for the embedded one) and L-stability; nothing here is taken from a binary.
#pragma once

#define SBZ_R5_GAMMA 0.21193756319429014

#define SBZ_R5_A21 3.0
#define SBZ_R5_A31 2.849394379747939
#define SBZ_R5_A32 0.45842242204463923
#define SBZ_R5_A41 (-6.954028509809101)
#define SBZ_R5_A42 2.489845061869568
#define SBZ_R5_A43 (-10.358996098473584)
#define SBZ_R5_A51 2.8029986275628964
#define SBZ_R5_A52 0.5072464736228206
#define SBZ_R5_A53 (-0.3988312541770524)
#define SBZ_R5_A54 (-0.04721187230404641)
#define SBZ_R5_A61 (-7.502846399306121)
#define SBZ_R5_A62 2.561846144803919
#define SBZ_R5_A63 (-11.627539656261098)
#define SBZ_R5_A64 (-0.18268767659942256)
#define SBZ_R5_A65 0.030198172008377946

#define SBZ_R5_C21 (-14.155112264123755)
#define SBZ_R5_C31 (-17.97296035885952)
#define SBZ_R5_C32 (-2.859693295451294)
#define SBZ_R5_C41 147.12150275711716
#define SBZ_R5_C42 (-1.41221402718213)
#define SBZ_R5_C43 71.68940251302358
#define SBZ_R5_C51 165.43517024871676
#define SBZ_R5_C52 (-0.4592823456491126)
#define SBZ_R5_C53 42.90938336958603
#define SBZ_R5_C54 (-5.961986721573306)
#define SBZ_R5_C61 24.854864614690072
#define SBZ_R5_C62 (-3.0009227002832186)
#define SBZ_R5_C63 47.4931110020768
#define SBZ_R5_C64 5.5814197821558125
#define SBZ_R5_C65 (-0.6610691825249471)
#define SBZ_R5_C71 30.91273214028599
#define SBZ_R5_C72 (-3.1208243349937974)
#define SBZ_R5_C73 77.79954646070892
#define SBZ_R5_C74 34.28646028294783
#define SBZ_R5_C75 (-19.097331116725623)
#define SBZ_R5_C76 (-28.087943162872662)
#define SBZ_R5_C81 37.80277123390563
#define SBZ_R5_C82 (-3.2571969029072276)
#define SBZ_R5_C83 112.26918849496327
#define SBZ_R5_C84 66.9347231244047
#define SBZ_R5_C85 (-40.06618937091002)
#define SBZ_R5_C86 (-54.66780262877968)
#define SBZ_R5_C87 (-9.48861652309627)

#define SBZ_R5_C2 0.6358126895828704
#define SBZ_R5_C3 0.4095798393397535
#define SBZ_R5_C4 0.9769306725060716
#define SBZ_R5_C5 0.4288403609558664

#define SBZ_R5_D1 0.21193756319429014
#define SBZ_R5_D2 (-0.42387512638858027)
#define SBZ_R5_D3 (-0.3384627126235924)
#define SBZ_R5_D4 1.8046452872882734
#define SBZ_R5_D5 2.325825639765069

// Step-size control: PI controller with OrdinaryDiffEqCore's defaults for an order-5 method
// (beta1 = 7/(10·5), beta2 = 2/(5·5), qmin = 1/5, qmax = 10, safety 9/10, qoldinit = 1e-4).
#define SBZ_CTL_BETA1 0.14
#define SBZ_CTL_BETA2 0.08
#define SBZ_CTL_QMIN 0.2
#define SBZ_CTL_QMAX 10.0
#define SBZ_CTL_GAMMA 0.9
#define SBZ_CTL_QOLDINIT 1e-4
