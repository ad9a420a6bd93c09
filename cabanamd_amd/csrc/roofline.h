// roofline.h -- work model of the NNP force step: FP32 flops per term of
// the symmetry functions (k_nnp_sf), the force chain (k_nnp_force) and the
// network layers, counted from the arithmetic the kernels write.
//
// Convention: an add, a multiply or a compare-free select counts 1 flop, a
// fused multiply-add 2, a transcendental or reciprocal (exp2, log2, sqrt,
// rcp, cos) 1; e^x is a multiply and an exp2 (2).  The counts are per term
// of the loops over
//   - neighbours j of atom i inside rc            (n_nb terms),
//   - (neighbour, radial (eta, rs) group)          (n_nb * nrg),
//   - neighbour pairs (j, k) with r_jk < rc        (triplets),
//   - (triplet, (lambda, zeta) pair)               (triplets * nap),
//   - (triplet, angular (eta, rs) group)           (triplets * nag),
//   - (triplet, angular function)                  (triplets * nag * nap),
// so the flops of a step are sum_k terms_k * F_k (nnp.hip counts the terms
// in k_nnp_sf).  Candidate pairs rejected by the r_jk test and the
// bookkeeping of the compacted pair stream are not counted: the model is
// the useful work, and the hardware counters (SQ_INSTS_VALU_FLOPS_FP32)
// measure what is issued.
#pragma once

namespace cmd
{
namespace nnp_flops
{
// k_nnp_sf
// neighbour: distance (3 sub + mul + 2 fma = 8), sqrt, rcp, unit vector
// (3 mul), cutoff 0.5 (cos(pi r / rc) + 1) (mul, cos, fma = 4) and its
// derivative (mul, sin-free form: 3)
constexpr double SF_NB = 20.0;
// radial group: t = r - rs (1), -eta t t (2), e^x (2), * fc (1), the
// element-matched accumulation (1)
constexpr double SF_RAD = 7.0;
// triplet: r_jk (3 sub + 5 + sqrt = 9), f1 f2 f3 with f3 = fcut(r_jk)
// (2 + 4), cos theta (dot 5 + 2 mul), 1 +- cos and its log2 for both signs
// of lambda (4)
constexpr double SF_TRIP = 26.0;
// (triplet, (lambda, zeta)): exp2(zeta log2 base) (mul + exp2), the
// 2^(1 - zeta) factor (mul), * F (mul)
constexpr double SF_ANGP = 4.0;
// (triplet, (eta, rs)): three differences (3), squares summed (5),
// * -eta (1), e^x (2)
constexpr double SF_ANGG = 11.0;
// (triplet, function): G += E_g P_p (fma)
constexpr double SF_ANG = 2.0;

// k_nnp_force
// neighbour: the gather above (20), s = gsum / r (1), s * u (3), the
// neighbour's and the atom's force (6)
constexpr double F_NB = 30.0;
// radial group: t (1), e^x with its argument (4), d = e (-2 eta t fc +
// dfc) (5), the coefficient sum over elements (1), gsum fma (2)
constexpr double F_RAD = 13.0;
// triplet: r_jk, 1/r_jk, f3 and df3 (9 + 1 + 4 + 4), F and its three
// partials (2 + 6), cos theta (7), d cos / d r for the three distances
// (12), the three force vectors from the radial and angular sums (18),
// the reactions kept in registers and LDS (9 + 6)
constexpr double F_TRIP = 78.0;
// (triplet, (lambda, zeta)): P and dP/dcos (exp2, 4 mul)
constexpr double F_ANGP = 6.0;
// (triplet, (eta, rs)): E_g as in the symmetry functions (11), its
// contribution to the four sums (8)
constexpr double F_ANGG = 19.0;
// (triplet, function): the coefficient-weighted P and dP sums (2 fma)
constexpr double F_ANG = 4.0;

// When every element's angular rs are 0 both kernels take a shorter path:
// the Gaussian argument r_ij^2 + r_ik^2 + r_jk^2 is formed once per triplet
// (5) and shared by every eta.
// k_nnp_sf: (triplet, (eta, rs)): * -eta (1), e^x (2), * F (1); the
// (lambda, zeta) factor loses its * F (3)
constexpr double SF_TRIP_RS0 = SF_TRIP + 5.0;
constexpr double SF_ANGP_RS0 = 3.0;
constexpr double SF_ANGG_RS0 = 4.0;
// k_nnp_force: (triplet, (eta, rs)): e^x with * -eta (3), w (1), G sum (1),
// dP sum (2), the shared gradient factor (2); per triplet that factor times
// the three distances (3)
constexpr double F_TRIP_RS0 = F_TRIP + 5.0 + 3.0;
constexpr double F_ANGG_RS0 = 9.0;

// network: forward 2 (nsf h1 + h1 h2 + h2) per atom, the backward pass
// (dE/dG through the same layers) twice that
constexpr double MLP_PASSES = 3.0;
} // namespace nnp_flops
} // namespace cmd
