// Host-side parameters of the triangle FMMs: leaf sizes and the quadrature
// rules on the reference triangle (see tria_rules.cpp).
#pragma once
#include <vector>

namespace fmm3d {

// Leaf size of the triangle FMMs (src/lfmm3dtria.f:297-304,
// src/hfmm3dtria.f:309-316): 8/3, 15/3, 30/3, 60/3, 120/3, 240/3, 480/3,
// 700/3 in integer arithmetic for iprec -2..5, every triangle and target in
// one leaf at iprec 6.
int tri_nbox(int iprec, int ns, int nt);

// The far-field rule of iprec: (xi, eta, w) triples on the reference
// triangle (xi, eta >= 0, xi + eta <= 1; weights summing to 1/2).
std::vector<double> tri_rule(int iprec, bool helmholtz);

// The D3-symmetric rule of degree 1..6 as (xi, eta, w) triples (weights
// summing to 1/2); empty for other degrees.
const std::vector<double>& tri_sym_rule(int degree);

// Degree of the near-field rule of the Helmholtz triangle FMM at iprec
// (src/hfmm3dtria.f:730-751): 2 below iprec 0, 4 at 0, 6 from 1 on.
inline int htri_nqtri(int iprec) { return iprec < 0 ? 2 : iprec == 0 ? 4 : 6; }

}  // namespace fmm3d
