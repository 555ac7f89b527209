// Host-side conversion of the reference's arrays (Fortran column-major,
// 1-based) into the device layout of the stage kernels, with the checks the
// C-ABI promises.  Shared by the single-shot and the batched contexts.
#pragma once
#include <cstring>
#include <string>
#include <vector>

namespace nexd {

struct OpsHost {
    std::vector<double> D;    // [3][NP][NP]
    std::vector<double> LT;   // [NFP][NP]
    std::vector<double> M;    // [NP][NP]
    std::vector<int> fm;      // [4][NPF] 0-based
};

inline std::string convert_operators(int np, int npf, const double *dr, const double *ds, const double *dt,
                                     const double *lift, const double *mass, const int *fmask, OpsHost &o) {
    if (!dr || !ds || !dt || !lift || !mass || !fmask) return "null operator";
    const int nfp = 4 * npf;
    o.D.assign((size_t)3 * np * np, 0.0);
    o.LT.assign((size_t)nfp * np, 0.0);
    o.M.assign((size_t)np * np, 0.0);
    o.fm.assign((size_t)4 * npf, 0);
    const double *src[3] = {dr, ds, dt};
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < np; ++i)
            for (int m = 0; m < np; ++m) o.D[((size_t)d * np + i) * np + m] = src[d][(size_t)m * np + i];
    for (int k = 0; k < nfp; ++k)
        for (int j = 0; j < np; ++j) o.LT[(size_t)k * np + j] = lift[(size_t)k * np + j];
    for (int i = 0; i < np; ++i)
        for (int k = 0; k < np; ++k) o.M[(size_t)i * np + k] = mass[(size_t)k * np + i];
    for (int f = 0; f < 4; ++f)
        for (int i = 0; i < npf; ++i) {
            const int v = fmask[f * npf + i] - 1;
            if (v < 0 || v >= np) return "fmask entry out of range";
            o.fm[f * npf + i] = v;
        }
    return "";
}

// Dr, Ds, Dt in the A-operand order of the 16x16x4 MFMA (row tile rt,
// k-step ks, lane l -> [rt][ks][l]); lane l of k-step ks holds tile row
// l%16 at column 4ks + l/16 (zero beyond Np).  Accumulator element v of a
// lane in lane group g = l/16 is tile row 4v + g (FP64 instruction) or
// 4g + v (FP32).  The rows are packed so that each lane group owns nodes
// g, g+4, g+8, ... with their three operator rows in order: element v of
// row tile rt is row q = 4rt + v of the group, node g + 4*(q/3), operator
// q%3.  A lane thus receives the three derivatives of each of its nodes,
// across consecutive tiles, with no padding row (3*ceil(Np/4) rows per
// group in ceil(3*ceil(Np/4)/4) tiles).
inline std::vector<double> mfma_operator_fragments(const std::vector<double> &D, int np, bool f64) {
    const int npg = (np + 3) / 4, nt = (3 * npg + 3) / 4, nk = (np + 3) / 4;
    std::vector<double> out((size_t)nt * nk * 64, 0.0);
    for (int rt = 0; rt < nt; ++rt)
        for (int ks = 0; ks < nk; ++ks)
            for (int l = 0; l < 64; ++l) {
                const int i = l % 16, m = 4 * ks + l / 16;
                const int g = f64 ? i % 4 : i / 4, v = f64 ? i / 4 : i % 4;
                const int q = 4 * rt + v, node = g + 4 * (q / 3), op = q % 3;
                if (node < np && m < np) out[((size_t)rt * nk + ks) * 64 + l] = D[((size_t)op * np + node) * np + m];
            }
    return out;
}

// [Dr | Ds | Dt] (Np x 3Np) in the A-operand order of the 16x16x4 MFMA, for
// the volume term of PML elements, which differentiates contravariant fluxes
// (each column needs sum_o D_o C_o): row tile rt holds nodes 16rt..16rt+15;
// k-step ks covers operator o = ks / KS at nodes 4(ks % KS)..+3 (KS =
// ceil(Np/4)); lane l holds node 16rt + l%16 at column 4(ks % KS) + l/16 of
// operator o (zero beyond Np).  Layout [rt][ks][l].
inline std::vector<double> mfma_operator_fragments_k(const std::vector<double> &D, int np) {
    const int nt = (np + 15) / 16, kso = (np + 3) / 4, nk = 3 * kso;
    std::vector<double> out((size_t)nt * nk * 64, 0.0);
    for (int rt = 0; rt < nt; ++rt)
        for (int ks = 0; ks < nk; ++ks)
            for (int l = 0; l < 64; ++l) {
                const int op = ks / kso, m = 4 * (ks % kso) + l / 16, node = 16 * rt + l % 16;
                if (node < np && m < np) out[((size_t)rt * nk + ks) * 64 + l] = D[((size_t)op * np + node) * np + m];
            }
    return out;
}

// LIFT in the A-operand order of the 16x16x4 MFMA: row tile mt holds nodes
// 16mt..16mt+15, k-step ks face nodes 4ks..4ks+3; lane l holds node
// 16mt + l%16 at face node 4ks + l/16 (zero beyond Np).  LT is [NFP][NP]
// with LT[k][j] = LIFT(j,k).
inline std::vector<double> mfma_lift_fragments(const std::vector<double> &LT, int np, int nfp) {
    const int nt = (np + 15) / 16, nk = nfp / 4;
    std::vector<double> out((size_t)nt * nk * 64, 0.0);
    for (int mt = 0; mt < nt; ++mt)
        for (int ks = 0; ks < nk; ++ks)
            for (int l = 0; l < 64; ++l) {
                const int i = 16 * mt + l % 16, k = 4 * ks + l / 16;
                if (i < np) out[((size_t)mt * nk + ks) * 64 + l] = LT[(size_t)k * np + i];
            }
    return out;
}

// The mass matrix in the A-operand order of the 16x16x4 MFMA (the energy's
// products M y): row tile mt holds nodes 16mt..16mt+15, k-step ks nodes
// 4ks..4ks+3; lane l holds M(16mt + l%16, 4ks + l/16), zero beyond Np.  M is
// [NP][NP] row-major.
inline std::vector<double> mfma_mass_fragments(const std::vector<double> &M, int np) {
    const int nt = (np + 15) / 16, nk = (np + 3) / 4;
    std::vector<double> out((size_t)nt * nk * 64, 0.0);
    for (int mt = 0; mt < nt; ++mt)
        for (int ks = 0; ks < nk; ++ks)
            for (int l = 0; l < 64; ++l) {
                const int i = 16 * mt + l % 16, k = 4 * ks + l / 16;
                if (i < np && k < np) out[((size_t)mt * nk + ks) * 64 + l] = M[(size_t)i * np + k];
            }
    return out;
}

struct MeshHost {
    std::vector<double> metric;  // [9][nglob]
    std::vector<double> jac;     // [nglob]
    std::vector<double> fscale;  // [nelem][NFP]
    std::vector<double> frame;   // [nelem][4][9]
    std::vector<int> nbr;        // [nelem][4][NPF]
    std::vector<int> ftype;      // [nelem][4]
};

// err is "" on success; unsupported features are reported with a leading '!'
inline std::string convert_mesh(int nelem, int np, int npf, const std::vector<int> &fm, const int *ibool,
                                const double *const metric_in[9], const double *jacobian, const double *fscale,
                                const int *neighbor, const int *face, const int *ibt, const int *ibn,
                                const double *nnormal, const double *stangential, const double *ttangential,
                                MeshHost &m) {
    for (int k = 0; k < 9; ++k)
        if (!metric_in[k]) return "null mesh array";
    if (!ibool || !jacobian || !fscale || !neighbor || !face || !ibt || !ibn || !nnormal || !stangential ||
        !ttangential)
        return "null mesh array";
    const size_t ng = (size_t)nelem * np;
    const int nfp = 4 * npf;
    for (size_t i = 0; i < ng; ++i)
        if (ibool[i] != (int)i + 1) return "!ibool must be (ie-1)*Np+j";
    m.metric.resize(9 * ng);
    for (int k = 0; k < 9; ++k) std::memcpy(m.metric.data() + k * ng, metric_in[k], sizeof(double) * ng);
    m.jac.assign(jacobian, jacobian + ng);
    m.fscale.assign(fscale, fscale + (size_t)nelem * nfp);
    m.frame.resize((size_t)nelem * 36);
    m.nbr.resize((size_t)nelem * 4 * npf);
    m.ftype.resize((size_t)nelem * 4);
    for (int e = 0; e < nelem; ++e) {
        for (int f = 0; f < 4; ++f) {
            for (int k = 0; k < 3; ++k) {
                m.frame[(e * 4 + f) * 9 + k] = nnormal[(e * 4 + f) * 3 + k];
                m.frame[(e * 4 + f) * 9 + 3 + k] = stangential[(e * 4 + f) * 3 + k];
                m.frame[(e * 4 + f) * 9 + 6 + k] = ttangential[(e * 4 + f) * 3 + k];
            }
            const int n = neighbor[e * 4 + f], fc = face[e * 4 + f];
            for (int i = 0; i < npf; ++i) {
                const size_t idx = ((size_t)e * 4 + f) * npf + i;
                if ((size_t)ibt[idx] != (size_t)e * np + fm[f * npf + i] + 1)
                    return "!ibt must be ibool(fmask(:,is),ie)";
            }
            if (n > 0) {
                if (n > nelem) return "neighbor out of range";
                if (fc < 1 || fc > 4) return "face of an interior face must be 1..4";
                m.ftype[e * 4 + f] = fc - 1;
                for (int i = 0; i < npf; ++i) {
                    const size_t idx = ((size_t)e * 4 + f) * npf + i;
                    const int v = ibn[idx] - 1;
                    if (v < 0 || (size_t)v >= ng) return "ibn out of range";
                    m.nbr[idx] = v;
                }
            } else if (n == 0) {
                if (fc != -1 && fc != -2) return "boundary face must be free (-1) or absorbing (-2)";
                m.ftype[e * 4 + f] = fc;
                for (int i = 0; i < npf; ++i) m.nbr[((size_t)e * 4 + f) * npf + i] = -1;
            } else {
                return "!partition interfaces (neighbor == -1) are not supported";
            }
        }
    }
    return "";
}

inline void pml_coefficients(size_t ng, const double *ddx, const double *ddy, const double *ddz,
                             const double *alphax, const double *alphay, const double *alphaz, const double *kx,
                             const double *ky, const double *kz, std::vector<double> &pc) {
    pc.resize(9 * ng);
    for (size_t g = 0; g < ng; ++g) {
        pc[0 * ng + g] = alphax[g]; pc[1 * ng + g] = alphay[g]; pc[2 * ng + g] = alphaz[g];
        pc[3 * ng + g] = ddx[g] / kx[g]; pc[4 * ng + g] = ddy[g] / ky[g]; pc[5 * ng + g] = ddz[g] / kz[g];
        pc[6 * ng + g] = kx[g]; pc[7 * ng + g] = ky[g]; pc[8 * ng + g] = kz[g];
    }
}

}  // namespace nexd
