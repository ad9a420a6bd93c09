// nnp.hip -- Behler-Parrinello neural-network potential: ForceNNP
// (force_types/force_nnp_cabana_neigh_impl.h:41-130) on one GPU.
//
// The reference runs n2p2's ModeCabana in three phases per force call
// (impl.h:93-98): symmetry functions, atomic networks, forces.  Here:
//   k_nnp_sf     one wave per atom: the atom's neighbours inside rc are
//                compacted into LDS (FP32 relative vectors from FP64
//                positions), radial functions with lanes over neighbours and
//                angular functions with lanes over neighbour pairs; per-lane
//                partial sums are combined with a reduce-scatter butterfly.
//   k_nnp_mlp    per element, 16 atoms per wave: the two hidden layers and
//                the backward pass as BF16 MFMA GEMMs (16x16x16, FP32
//                accumulate) over the scaled symmetry functions; output
//                layer, tanh and scaling on the VALU.  k_nnp_mlp_f32 is the
//                same network in FP32 (one atom per lane).
//   k_nnp_force  one wave per atom: dE/dG chained into dG/dr on the same
//                pairs; reactions on neighbours gather in LDS and reach
//                f by FP64 atomics (ghost rows included, folded by
//                update_force as in the reference, cabanamd_impl.h:346-353).
// Definitions (n2p2's, for the input.nn keywords used by input/nnp):
// cosine cutoff fc(r) = (cos(pi r/rc) + 1)/2; radial G = sum_j
// exp(-eta (r-rs)^2) fc; angular G = 2^(1-zeta) sum_{j<k} (1 + lambda cos)^zeta
// exp(-eta sum (r-rs)^2) fc fc fc; min-max scaling; tanh/linear layers.
//
// The symmetry-function set of an element must be a grid: radial functions
// = {(eta, rs)} x {neighbour element or any}, angular = {(eta, rs)} x
// {(lambda, zeta)} over any neighbour pair.  Each pair then needs one exp
// per (eta, rs) and one power per (lambda, zeta), and every term is a
// product of the two.  Both the reference fixture and the benchmark model
// have this form; other sets are rejected at setup.
#include "internal.h"
#include "roofline.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

// The NNP kernels compute in FP32 against a tolerance, not bit for bit:
// multiply-adds may fuse here (the library builds with contraction off for
// the FP64 paths, whose sums must match the oracle exactly).
#pragma clang fp contract( fast )

namespace cmd
{
static constexpr int NNP_MAXEL = 4;
static constexpr int RG_MAX = 16; // radial (eta, rs) values
static constexpr int RE_MAX = 2;  // radial neighbour-element filters
static constexpr int AG_MAX = 8;  // angular (eta, rs) values
static constexpr int AP_MAX = 8;  // angular (lambda, zeta) values
static constexpr int NR_MAX = RG_MAX * RE_MAX; // 32
static constexpr int NA_MAX = AG_MAX * AP_MAX; // 64
static constexpr int NSF_LD = 96;              // row stride of G / dE/dG
static constexpr int KPAD = 96;                // MFMA K of layer 1 (6 x 16)
static constexpr int HPAD = 32;                // hidden width (2 x 16)
static constexpr int NN_MAX = 128;             // neighbours inside rc
static constexpr int TCAP = 512;               // compacted neighbour pairs per wave
static constexpr int NNP_TPB = 256;

struct NnpElemPlan
{
    int nsf;
    int nrg, nre, nag, nap;
    float r_eta[RG_MAX], r_rs[RG_MAX];
    int r_el[RE_MAX];              // neighbour element, -1 = any
    int r_out[RG_MAX][RE_MAX];     // model index of radial (g, e)
    float a_eta[AG_MAX], a_rs[AG_MAX];
    float a_lam[AP_MAX], a_zeta[AP_MAX], a_lpref[AP_MAX]; // lpref = 1 - zeta
    int a_out[AG_MAX][AP_MAX];     // model index of angular (g, p)
    float g_mul[NSF_LD], g_add[NSF_LD]; // Gs = g_mul * G + g_add
    int h1, h2;
    int act1, act2, act3; // 1 = tanh, 0 = linear
    float W1[HPAD][KPAD]; // [out][in], zero padded
    float b1[HPAD];
    float W2[HPAD][HPAD];
    float b2[HPAD];
    float w3[HPAD];
    float b3;
    // The layer weights as three BF16 parts (w = hi + mid + lo to FP32
    // precision), laid out as the MFMA B operands read them: [out][in] for
    // the forward products, [in][out] for the backward ones.
    alignas( 16 ) unsigned short W1s[3][HPAD][KPAD];
    alignas( 16 ) unsigned short W1Ts[3][KPAD][HPAD];
    alignas( 16 ) unsigned short W2s[3][HPAD][HPAD];
    alignas( 16 ) unsigned short W2Ts[3][HPAD][HPAD];
};

struct NnpPlan
{
    int nelem;
    float rc, rc2, pi_rc; // pi_rc = pi / rc
    NnpElemPlan el[NNP_MAXEL];
};
} // namespace cmd

// Host-side model: what input.nn, scaling.data and weights.NNN.data hold.
struct cmd_nnp_model
{
    struct SF
    {
        int type, e1, e2;
        double eta, rs, lambda, zeta, rc;
    };
    struct Elem
    {
        std::string name;
        int z = 0;
        std::vector<SF> sf;
        std::vector<double> gmin, gmax;
        std::vector<std::vector<double>> W, b; // per layer [out][in], [out]
        double atom_energy = 0.0;
    };
    int nelem = 0;
    int nlayers = 3;
    std::vector<int> hidden;
    std::string act = "ttl";
    bool scale = true;
    double smin = 0.0, smax = 1.0;
    std::vector<Elem> el;
};

struct cmd_nnp
{
    cmd_nnp_model model;
    cmd::NnpPlan plan{};
    cmd::DevBuf<cmd::NnpPlan> dplan;
    bool plan_ok = false;
    int prec = CMD_PREC_MIXED; // MIXED: BF16 MFMA layers; FP64: FP32 layers
    cmd::DevBuf<float> G, dEdG, Ei;
    cmd::DevBuf<int> elist, ecount, flag;
    cmd::DevBuf<double> part;
    cmd::DevBuf<int4> terms; // per atom: neighbours, triplets, element (last compute)
    int cap = 0;
    int n_last = 0;
};

namespace cmd
{
// ---------------------------------------------------------------- plan
static int find_or_add( std::vector<std::pair<double, double>> &v, double a,
                        double b )
{
    for ( size_t k = 0; k < v.size(); ++k )
        if ( v[k].first == a && v[k].second == b )
            return (int)k;
    v.push_back( { a, b } );
    return (int)v.size() - 1;
}

// BF16 bits of a float, rounded to nearest even (the format's NaN and
// infinity are not produced from the finite weights and activations here).
__host__ __device__ inline unsigned short bf16_bits( float f )
{
    unsigned u;
    std::memcpy( &u, &f, sizeof( u ) );
    u += 0x7fffu + ( ( u >> 16 ) & 1u );
    return (unsigned short)( u >> 16 );
}

__host__ __device__ inline float bf16_value( unsigned short b )
{
    const unsigned u = (unsigned)b << 16;
    float f;
    std::memcpy( &f, &u, sizeof( f ) );
    return f;
}

// x = hi + mid + lo: each part the BF16 rounding of what the parts before it
// leave, so the three hold the 24 significant bits of an FP32 value.
__host__ __device__ inline void bf16_split3( float x, unsigned short part[3] )
{
    part[0] = bf16_bits( x );
    const float r1 = x - bf16_value( part[0] );
    part[1] = bf16_bits( r1 );
    const float r2 = r1 - bf16_value( part[1] );
    part[2] = bf16_bits( r2 );
}

static cmd_status build_plan( const cmd_nnp_model &m, NnpPlan &P )
{
    std::memset( &P, 0, sizeof( P ) );
    CMD_REQUIRE( m.nelem >= 1 && m.nelem <= NNP_MAXEL, CMD_ERR_ARG,
                 "NNP models with 1..%d elements are supported", NNP_MAXEL );
    CMD_REQUIRE( m.hidden.size() == 2, CMD_ERR_ARG,
                 "NNP networks with two hidden layers are supported" );
    CMD_REQUIRE( m.hidden[0] >= 1 && m.hidden[0] <= HPAD && m.hidden[1] >= 1 &&
                     m.hidden[1] <= HPAD,
                 CMD_ERR_ARG, "hidden layers wider than %d nodes", HPAD );
    CMD_REQUIRE( m.act.size() == 3 && m.act[2] == 'l', CMD_ERR_ARG,
                 "activations must be t/l for the hidden layers and l for "
                 "the output" );
    P.nelem = m.nelem;
    double rc = -1.0;
    for ( int e = 0; e < m.nelem; ++e )
    {
        const auto &E = m.el[e];
        NnpElemPlan &Q = P.el[e];
        const int ns = (int)E.sf.size();
        CMD_REQUIRE( ns >= 1 && ns <= NSF_LD && ns <= KPAD, CMD_ERR_ARG,
                     "element %d has %d symmetry functions (1..%d supported)",
                     e, ns, KPAD );
        Q.nsf = ns;
        std::vector<std::pair<double, double>> rg, ag, ap;
        std::vector<int> re;
        for ( const auto &s : E.sf )
        {
            if ( rc < 0 )
                rc = s.rc;
            CMD_REQUIRE( s.rc == rc, CMD_ERR_ARG,
                         "all symmetry functions must share one cutoff" );
            if ( s.type == 2 )
            {
                find_or_add( rg, s.eta, s.rs );
                const int ne = m.nelem == 1 ? -1 : s.e1;
                if ( std::find( re.begin(), re.end(), ne ) == re.end() )
                    re.push_back( ne );
            }
            else if ( s.type == 3 )
            {
                CMD_REQUIRE( m.nelem == 1 || s.e1 < 0, CMD_ERR_ARG,
                             "angular functions must accept any neighbour "
                             "pair" );
                find_or_add( ag, s.eta, s.rs );
                find_or_add( ap, s.lambda, s.zeta );
            }
            else
            {
                set_error( "symmetry function type %d is not supported", s.type );
                return CMD_ERR_ARG;
            }
        }
        CMD_REQUIRE( (int)rg.size() <= RG_MAX && (int)re.size() <= RE_MAX &&
                         (int)ag.size() <= AG_MAX && (int)ap.size() <= AP_MAX,
                     CMD_ERR_ARG, "symmetry-function grid too large" );
        Q.nrg = (int)rg.size();
        Q.nre = (int)re.size();
        Q.nag = (int)ag.size();
        Q.nap = (int)ap.size();
        for ( int g = 0; g < RG_MAX; ++g )
            for ( int k = 0; k < RE_MAX; ++k )
                Q.r_out[g][k] = -1;
        for ( int g = 0; g < AG_MAX; ++g )
            for ( int k = 0; k < AP_MAX; ++k )
                Q.a_out[g][k] = -1;
        for ( int g = 0; g < Q.nrg; ++g )
        {
            Q.r_eta[g] = (float)rg[g].first;
            Q.r_rs[g] = (float)rg[g].second;
        }
        for ( int k = 0; k < Q.nre; ++k )
            Q.r_el[k] = re[k];
        for ( int g = 0; g < Q.nag; ++g )
        {
            Q.a_eta[g] = (float)ag[g].first;
            Q.a_rs[g] = (float)ag[g].second;
        }
        for ( int p = 0; p < Q.nap; ++p )
        {
            Q.a_lam[p] = (float)ap[p].first;
            Q.a_zeta[p] = (float)ap[p].second;
            Q.a_lpref[p] = (float)( 1.0 - ap[p].second );
        }
        for ( int s = 0; s < ns; ++s )
        {
            const auto &f = E.sf[s];
            if ( f.type == 2 )
            {
                const int g = find_or_add( rg, f.eta, f.rs );
                const int ne = m.nelem == 1 ? -1 : f.e1;
                const int k = (int)( std::find( re.begin(), re.end(), ne ) - re.begin() );
                CMD_REQUIRE( Q.r_out[g][k] < 0, CMD_ERR_ARG,
                             "duplicate radial symmetry function" );
                Q.r_out[g][k] = s;
            }
            else
            {
                const int g = find_or_add( ag, f.eta, f.rs );
                const int p = find_or_add( ap, f.lambda, f.zeta );
                CMD_REQUIRE( Q.a_out[g][p] < 0, CMD_ERR_ARG,
                             "duplicate angular symmetry function" );
                Q.a_out[g][p] = s;
            }
        }
        for ( int g = 0; g < Q.nrg; ++g )
            for ( int k = 0; k < Q.nre; ++k )
                CMD_REQUIRE( Q.r_out[g][k] >= 0, CMD_ERR_ARG,
                             "radial functions do not form a grid" );
        for ( int g = 0; g < Q.nag; ++g )
            for ( int p = 0; p < Q.nap; ++p )
                CMD_REQUIRE( Q.a_out[g][p] >= 0, CMD_ERR_ARG,
                             "angular functions do not form a grid" );
        for ( int s = 0; s < NSF_LD; ++s )
        {
            Q.g_mul[s] = 0.0f;
            Q.g_add[s] = 0.0f;
        }
        for ( int s = 0; s < ns; ++s )
        {
            if ( m.scale )
            {
                const double w = E.gmax[s] - E.gmin[s];
                const double mul = w > 0.0 ? ( m.smax - m.smin ) / w : 0.0;
                Q.g_mul[s] = (float)mul;
                Q.g_add[s] = (float)( m.smin - mul * E.gmin[s] );
            }
            else
                Q.g_mul[s] = 1.0f;
        }
        Q.h1 = m.hidden[0];
        Q.h2 = m.hidden[1];
        Q.act1 = m.act[0] == 't';
        Q.act2 = m.act[1] == 't';
        Q.act3 = 0;
        CMD_REQUIRE( E.W.size() == 3 && E.W[0].size() == (size_t)Q.h1 * ns &&
                         E.W[1].size() == (size_t)Q.h2 * Q.h1 &&
                         E.W[2].size() == (size_t)Q.h2,
                     CMD_ERR_ARG, "network weights of element %d are missing", e );
        for ( int o = 0; o < Q.h1; ++o )
        {
            for ( int k = 0; k < ns; ++k )
                Q.W1[o][k] = (float)E.W[0][(size_t)o * ns + k];
            Q.b1[o] = (float)E.b[0][o];
        }
        for ( int o = 0; o < Q.h2; ++o )
        {
            for ( int k = 0; k < Q.h1; ++k )
                Q.W2[o][k] = (float)E.W[1][(size_t)o * Q.h1 + k];
            Q.b2[o] = (float)E.b[1][o];
        }
        for ( int k = 0; k < Q.h2; ++k )
            Q.w3[k] = (float)E.W[2][k];
        Q.b3 = (float)E.b[2][0];
        for ( int o = 0; o < HPAD; ++o )
        {
            for ( int k = 0; k < KPAD; ++k )
            {
                unsigned short part[3];
                bf16_split3( Q.W1[o][k], part );
                for ( int q = 0; q < 3; ++q )
                {
                    Q.W1s[q][o][k] = part[q];
                    Q.W1Ts[q][k][o] = part[q];
                }
            }
            for ( int k = 0; k < HPAD; ++k )
            {
                unsigned short part[3];
                bf16_split3( Q.W2[o][k], part );
                for ( int q = 0; q < 3; ++q )
                {
                    Q.W2s[q][o][k] = part[q];
                    Q.W2Ts[q][k][o] = part[q];
                }
            }
        }
    }
    P.rc = (float)rc;
    P.rc2 = (float)( rc * rc );
    P.pi_rc = (float)( 3.14159265358979323846 / rc );
    return CMD_OK;
}

// ---------------------------------------------------------------- device
// Cosine cutoff 0.5 (cos(pi r / rc) + 1) and its derivative; k = pi / rc.
__device__ __forceinline__ float fcut( float r, float rc, float k )
{
    return r < rc ? 0.5f * ( __cosf( k * r ) + 1.0f ) : 0.0f;
}
__device__ __forceinline__ float dfcut( float r, float rc, float k )
{
    return r < rc ? -0.5f * k * __sinf( k * r ) : 0.0f;
}

// The angular constants of one element, held in scalar registers through
// the triplet loops (read once per wave from LDS, so the compiler cannot
// reload them from memory inside the loop).
struct AngK
{
    int nag, nap;
    unsigned lampos; // bit pp: lambda_pp > 0
    float eta[AG_MAX], rs[AG_MAX];
    float zeta[AP_MAX], lpref[AP_MAX], zm1[AP_MAX], zl[AP_MAX];
};

__device__ __forceinline__ AngK load_angk( const NnpElemPlan &Q, float *stage )
{
    const int lane = threadIdx.x & 63;
    // stage: eta[8] rs[8] zeta[8] lpref[8] lam[8]
    if ( lane < 40 )
    {
        const int k = lane & 7, w = lane >> 3;
        float v = 0.f;
        if ( w == 0 )
            v = Q.a_eta[k];
        else if ( w == 1 )
            v = Q.a_rs[k];
        else if ( w == 2 )
            v = Q.a_zeta[k];
        else if ( w == 3 )
            v = Q.a_lpref[k];
        else
            v = Q.a_lam[k];
        stage[lane] = v;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    AngK K;
    K.nag = __builtin_amdgcn_readfirstlane( Q.nag );
    K.nap = __builtin_amdgcn_readfirstlane( Q.nap );
    K.lampos = 0u;
#pragma unroll
    for ( int k = 0; k < AG_MAX; ++k )
    {
        K.eta[k] = __builtin_bit_cast( float, __builtin_amdgcn_readfirstlane(
                                                   __builtin_bit_cast( int, stage[k] ) ) );
        K.rs[k] = __builtin_bit_cast( float, __builtin_amdgcn_readfirstlane(
                                                  __builtin_bit_cast( int, stage[8 + k] ) ) );
    }
#pragma unroll
    for ( int k = 0; k < AP_MAX; ++k )
    {
        const float z = __builtin_bit_cast( float, __builtin_amdgcn_readfirstlane(
                                                       __builtin_bit_cast( int, stage[16 + k] ) ) );
        const float lp = __builtin_bit_cast( float, __builtin_amdgcn_readfirstlane(
                                                        __builtin_bit_cast( int, stage[24 + k] ) ) );
        const float lam = __builtin_bit_cast( float, __builtin_amdgcn_readfirstlane(
                                                         __builtin_bit_cast( int, stage[32 + k] ) ) );
        K.zeta[k] = z;
        K.lpref[k] = lp;
        K.zm1[k] = z - 1.f;
        K.zl[k] = z * lam;
        if ( lam > 0.f )
            K.lampos |= 1u << k;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    return K;
}

// Angular factors of the element's (lambda, zeta) grid for cos theta = cs:
// Pw = 2^(1-zeta) (1 + lambda cs)^zeta through the hardware log2/exp2
// (exp2(-inf) = 0), and Dw = zeta lambda 2^(1-zeta) (1 + lambda cs)^(zeta-1),
// the derivative of Pw with respect to cs, as zeta lambda Pw / base.  For
// zeta = 1, Dw = lambda, also at base = 0 (collinear neighbours); for
// zeta > 1 at base = 0 both are 0 (the reciprocal is taken as 0 there).
template <bool WANT_D>
__device__ __forceinline__ void ang_powers( const AngK &K, float cs, float *Pw, float *Dw )
{
    const float bp = fmaxf( 1.0f + cs, 0.0f ), bm = fmaxf( 1.0f - cs, 0.0f );
    const float lbp = __builtin_amdgcn_logf( bp );
    const float lbm = __builtin_amdgcn_logf( bm );
    float ibp = 0.0f, ibm = 0.0f;
    if ( WANT_D )
    {
        ibp = bp > 0.0f ? __builtin_amdgcn_rcpf( bp ) : 0.0f;
        ibm = bm > 0.0f ? __builtin_amdgcn_rcpf( bm ) : 0.0f;
    }
#pragma unroll
    for ( int pp = 0; pp < AP_MAX; ++pp )
    {
        Pw[pp] = 0.0f;
        if ( WANT_D )
            Dw[pp] = 0.0f;
        if ( pp < K.nap )
        {
            const bool pos = ( K.lampos >> pp ) & 1u;
            const float lb = pos ? lbp : lbm;
            Pw[pp] = __builtin_amdgcn_exp2f( K.zeta[pp] * lb + K.lpref[pp] );
            if ( WANT_D )
                Dw[pp] = K.zm1[pp] != 0.f ? K.zl[pp] * Pw[pp] * ( pos ? ibp : ibm ) : K.zl[pp];
        }
    }
}

// Reduce-scatter across the 64 lanes: in: v[S] per lane; out: v[0] of
// lane l = the sum over lanes of value l (S = 64) or l & 31 (S = 32).
// Each halving step keeps the half selected by one lane bit and sends the
// other half to the partner lane, so S values cost S - 1 shuffles in all.
template <int M>
__device__ __forceinline__ void rs_step( float *v, int lane )
{
    if constexpr ( M > 1 )
    {
        constexpr int O = M / 2;
        const bool hi = lane & O;
#pragma unroll
        for ( int k = 0; k < O; ++k )
        {
            const float keep = hi ? v[O + k] : v[k];
            const float send = hi ? v[k] : v[O + k];
            v[k] = keep + __shfl_xor( send, O, 64 );
        }
        rs_step<O>( v, lane );
    }
}

template <int S>
__device__ __forceinline__ void wave_reduce_scatter( float *v )
{
    const int lane = threadIdx.x & 63;
    rs_step<S>( v, lane );
    if constexpr ( S == 32 )
        v[0] += __shfl_xor( v[0], 32, 64 );
}

// Neighbours of one atom inside rc: d = (x_j - x_i, r) and
// e = (1 / r, fc(r), dfc(r), element as float bits), two 16-byte reads each.
struct NbLds
{
    float4 d[NN_MAX];
    float4 e[NN_MAX];
    int j[NN_MAX];
};

// Compact the neighbours of atom i inside rc into LDS; returns the count
// (wave-uniform).  Sets *flag on overflow.
__device__ int gather_neighbours( int i, const double4 *__restrict__ x,
                                  const int64_t *__restrict__ off,
                                  const int *__restrict__ idx, float rc,
                                  float rc2, float kpi, NbLds &L,
                                  int *__restrict__ flag )
{
    const int lane = threadIdx.x & 63;
    const double4 xi = x[i];
    const int64_t o0 = off[i];
    const int n = (int)( off[i + 1] - o0 );
    int cnt = 0;
    for ( int base = 0; base < n; base += 64 )
    {
        const int k = base + lane;
        bool in = false;
        float dx = 0.f, dy = 0.f, dz = 0.f, r2 = 0.f;
        int j = 0;
        if ( k < n )
        {
            j = idx[o0 + k];
            const double4 xj = x[j];
            dx = (float)( xj.x - xi.x );
            dy = (float)( xj.y - xi.y );
            dz = (float)( xj.z - xi.z );
            r2 = dx * dx + dy * dy + dz * dz;
            in = r2 < rc2;
        }
        const unsigned long long m = __ballot( in );
        const int pos = cnt + (int)__builtin_amdgcn_mbcnt_hi(
                                  (unsigned)( m >> 32 ),
                                  __builtin_amdgcn_mbcnt_lo( (unsigned)m, 0u ) );
        if ( in && pos < NN_MAX )
        {
            const float r = __builtin_amdgcn_sqrtf( r2 );
            L.d[pos] = make_float4( dx, dy, dz, r );
            L.e[pos] = make_float4( __builtin_amdgcn_rcpf( r ), fcut( r, rc, kpi ),
                                    dfcut( r, rc, kpi ),
                                    __builtin_bit_cast( float, (int)x[j].w ) );
            L.j[pos] = j;
        }
        cnt += (int)__popcll( m );
    }
    if ( cnt > NN_MAX )
    {
        if ( lane == 0 )
            atomicOr( flag, 1 );
        cnt = NN_MAX;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    return cnt;
}

// Pair index p of n items -> (a, q), a < q, row-major over a.
__device__ __forceinline__ void pair_of( int p, int n, int &a, int &q )
{
    const float tn = 2.0f * n - 1.0f;
    int aa = (int)( ( tn - sqrtf( tn * tn - 8.0f * (float)p ) ) * 0.5f );
    aa = max( 0, min( aa, n - 2 ) );
    // start of row aa: aa*(2n - aa - 1)/2
    while ( aa > 0 && aa * ( 2 * n - aa - 1 ) / 2 > p )
        --aa;
    while ( ( aa + 1 ) * ( 2 * n - aa - 2 ) / 2 <= p )
        ++aa;
    a = aa;
    q = p - aa * ( 2 * n - aa - 1 ) / 2 + aa + 1;
}

// The neighbour pairs (a < q) of the gathered list whose third side is
// inside the cutoff, |x_q - x_a| < rc: tested 64 at a time and compacted
// into buf (TCAP entries per wave, a | q << 16), then handed to body(a, q)
// with every lane busy (lane l takes entries l*per .. l*per + per - 1).
// Used where the body keeps many accumulators in registers.
// With PAIRS the body takes two pairs at once, (a0, q0) and (a1, q1), the
// second valid when ok1 (else a repeat of the first): the two gradient
// chains are independent, so their latencies overlap.
template <bool PAIRS = false, class Body>
__device__ __forceinline__ void for_triplets( const NbLds &L, int n, float rc2,
                                             unsigned *buf, Body body, int *count = nullptr )
{
    const int lane = threadIdx.x & 63;
    // candidate pairs, 64 per test: with n <= 64 rows a and n - 2 - a side
    // by side (n - 1 - a and a + 1 entries: n in all), else the flat pair
    // index through pair_of
    const bool rows = n <= 64;
    const int np = n * ( n - 1 ) / 2;
    const int nit = rows ? n / 2 : ( np + 63 ) / 64; // rows: pairs of rows
    for ( int it = 0; it < nit; )
    {
        int cnt = 0;
        while ( cnt <= TCAP - 64 && it < nit )
        {
            bool ok = false;
            unsigned code = 0u;
            int a = 0, q = 0;
            bool have;
            if ( rows )
            {
                // rows a1 < a2 hold n - 1 - a1 and n - 1 - a2 = a1 + 1
                // entries; the middle row of an even n comes alone
                const int a1 = it, a2 = n - 2 - it;
                const int len1 = n - 1 - a1, len2 = a2 > a1 ? n - 1 - a2 : 0;
                have = lane < len1 + len2;
                if ( lane < len1 )
                {
                    a = a1;
                    q = a1 + 1 + lane;
                }
                else
                {
                    a = a2;
                    q = a2 + 1 + ( lane - len1 );
                }
            }
            else
            {
                const int p = it * 64 + lane;
                have = p < np;
                if ( have )
                    pair_of( p, n, a, q );
            }
            if ( have )
            {
                const float4 A = L.d[a], B = L.d[q];
                const float cx = B.x - A.x, cy = B.y - A.y, cz = B.z - A.z;
                ok = cx * cx + cy * cy + cz * cz < rc2;
                code = (unsigned)a | ( (unsigned)q << 16 );
            }
            const unsigned long long m = __ballot( ok );
            if ( ok )
                buf[cnt + (int)__builtin_amdgcn_mbcnt_hi(
                              (unsigned)( m >> 32 ),
                              __builtin_amdgcn_mbcnt_lo( (unsigned)m, 0u ) )] = code;
            cnt += (int)__popcll( m );
            ++it;
        }
        __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
        if ( count )
            *count += cnt;
        const int per = ( cnt + 63 ) >> 6;
        if constexpr ( PAIRS )
        {
            for ( int j = 0; j < per; j += 2 )
            {
                const int e = lane * per + j;
                if ( e < cnt )
                {
                    const unsigned c0 = buf[e];
                    const bool ok1 = j + 1 < per && e + 1 < cnt;
                    const unsigned c1 = ok1 ? buf[e + 1] : c0;
                    body( (int)( c0 & 0xffffu ), (int)( c0 >> 16 ), (int)( c1 & 0xffffu ),
                          (int)( c1 >> 16 ), ok1 );
                }
            }
        }
        else
        {
            for ( int j = 0; j < per; ++j )
            {
                const int e = lane * per + j;
                if ( e < cnt )
                {
                    const unsigned c = buf[e];
                    body( (int)( c & 0xffffu ), (int)( c >> 16 ) );
                }
            }
        }
        __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    }
}

// NAG, NAP > 0: the angular (eta, rs) and (lambda, zeta) grid sizes of
// every element, fixed at compile time (loops unrolled straight); 0: read
// from the plan.  RS0: every element's angular rs are 0, so the Gaussian
// argument r_ij^2 + r_ik^2 + r_jk^2 is shared by all eta (a kernel of its
// own: with both paths in one kernel the registers of the longer one would
// cost occupancy).
template <int NAG, int NAP, bool RS0>
__global__ __launch_bounds__( NNP_TPB ) void k_nnp_sf(
    int n_local, const double4 *__restrict__ x, const int64_t *__restrict__ off,
    const int *__restrict__ idx, const NnpPlan *__restrict__ P,
    float *__restrict__ G, int *__restrict__ flag, int4 *__restrict__ terms )
{
    __shared__ NbLds lds[NNP_TPB / 64];
    __shared__ unsigned tbuf[NNP_TPB / 64][TCAP];
    __shared__ float kst[NNP_TPB / 64][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * ( NNP_TPB / 64 ) + wv;
    if ( i >= n_local )
        return;
    NbLds &L = lds[wv];
    const float rc = P->rc, rc2 = P->rc2, kpi = P->pi_rc;
    const int n = gather_neighbours( i, x, off, idx, rc, rc2, kpi, L, flag );
    const int e = __builtin_amdgcn_readfirstlane( (int)x[i].w );
    const NnpElemPlan &Q = P->el[e];
    AngK K = load_angk( Q, kst[wv] );
    if constexpr ( NAG > 0 )
        K.nag = NAG;
    if constexpr ( NAP > 0 )
        K.nap = NAP;
    float *Gi = G + (size_t)i * NSF_LD;
    // radial: lanes over neighbours
    {
        float acc[NR_MAX];
#pragma unroll
        for ( int s = 0; s < NR_MAX; ++s )
            acc[s] = 0.0f;
        for ( int a = lane; a < n; a += 64 )
        {
            const float4 Ea = L.e[a];
            const float r = L.d[a].w, fc = Ea.y;
            const int ej = __builtin_bit_cast( int, Ea.w );
#pragma unroll
            for ( int g = 0; g < RG_MAX; ++g )
            {
                if ( g < Q.nrg )
                {
                    const float t = r - Q.r_rs[g];
                    const float v = __expf( -Q.r_eta[g] * t * t ) * fc;
#pragma unroll
                    for ( int k = 0; k < RE_MAX; ++k )
                        if ( k < Q.nre && ( Q.r_el[k] < 0 || Q.r_el[k] == ej ) )
                            acc[g * RE_MAX + k] += v;
                }
            }
        }
        wave_reduce_scatter<NR_MAX>( acc );
        if ( lane < NR_MAX )
        {
            const int g = lane / RE_MAX, k = lane % RE_MAX;
            if ( g < Q.nrg && k < Q.nre )
                Gi[Q.r_out[g][k]] = acc[0];
        }
    }
    // angular: lanes over the compacted neighbour pairs
    {
        float acc[NA_MAX];
#pragma unroll
        for ( int s = 0; s < NA_MAX; ++s )
            acc[s] = 0.0f;
        int ntrip = 0;
        for_triplets(
            L, n, rc2, tbuf[wv], [&]( int a, int q ) {
            const float4 A = L.d[a], B = L.d[q], Ea = L.e[a], Eb = L.e[q];
            const float cx = B.x - A.x, cy = B.y - A.y, cz = B.z - A.z;
            const float r3 = __builtin_amdgcn_sqrtf( cx * cx + cy * cy + cz * cz );
            const float F = Ea.y * Eb.y * fcut( r3, rc, kpi );
            const float cs = ( A.x * B.x + A.y * B.y + A.z * B.z ) * ( Ea.x * Eb.x );
            float Pw[AP_MAX];
            ang_powers<false>( K, cs, Pw, nullptr );
            if constexpr ( RS0 )
            {
                // one Gaussian argument, and the cutoff product joins the
                // Gaussian (nag multiplies, not nap)
                const float s2 = A.w * A.w + B.w * B.w + r3 * r3;
#pragma unroll
                for ( int g = 0; g < AG_MAX; ++g )
                {
                    if ( g < K.nag )
                    {
                        const float EgF = __expf( -K.eta[g] * s2 ) * F;
#pragma unroll
                        for ( int pp = 0; pp < AP_MAX; ++pp )
                            acc[g * AP_MAX + pp] += EgF * Pw[pp];
                    }
                }
            }
            else
            {
#pragma unroll
                for ( int pp = 0; pp < AP_MAX; ++pp )
                    Pw[pp] *= F;
#pragma unroll
                for ( int g = 0; g < AG_MAX; ++g )
                {
                    if ( g < K.nag )
                    {
                        const float rs = K.rs[g];
                        const float t1 = A.w - rs, t2 = B.w - rs, t3 = r3 - rs;
                        const float Eg = __expf( -K.eta[g] * ( t1 * t1 + t2 * t2 + t3 * t3 ) );
#pragma unroll
                        for ( int pp = 0; pp < AP_MAX; ++pp )
                            acc[g * AP_MAX + pp] += Eg * Pw[pp];
                    }
                }
            }
            },
            &ntrip );
        // the work model's term counts (roofline.h) of this atom: its
        // neighbours and triplets, and its element for the per-element
        // grids (one store per atom, summed when asked for)
        if ( terms && lane == 0 )
            terms[i] = make_int4( n, ntrip, e, 0 );
        wave_reduce_scatter<NA_MAX>( acc );
        const int g = lane / AP_MAX, pp = lane % AP_MAX;
        if ( g < Q.nag && pp < Q.nap )
            Gi[Q.a_out[g][pp]] = acc[0];
    }
}

// ---- network: split-BF16 MFMA, 16 atoms of one element per wave
//
// Every layer product runs on gfx950's 16x16x32 BF16 MFMA with FP32
// accumulation.  Each FP32 operand is split into three BF16 parts (a = ah +
// am + al, exact for FP32 values), and a product sums the six part products
// of order up to 2^-16 relative (ah bh, ah bm, am bh, am bm, ah bl, al bh),
// smallest first: the result matches an FP32 product to about 2^-24, while a
// single BF16 product carries 2^-8.  The weights come pre-split from the
// plan; activations are split as they are loaded.
typedef __bf16 bf16x8 __attribute__( ( ext_vector_type( 8 ) ) );
typedef short s16x8 __attribute__( ( ext_vector_type( 8 ) ) );
typedef float f32x4 __attribute__( ( ext_vector_type( 4 ) ) );

__device__ __forceinline__ float act_f( float z, int tanh_ )
{
    return tanh_ ? tanhf( z ) : z;
}

// The three BF16 parts of 8 consecutive FP32 values.
__device__ __forceinline__ void split8( const float *v, bf16x8 &h, bf16x8 &m, bf16x8 &l )
{
    s16x8 sh, sm, sl;
#pragma unroll
    for ( int u = 0; u < 8; ++u )
    {
        unsigned short part[3];
        bf16_split3( v[u], part );
        sh[u] = (short)part[0];
        sm[u] = (short)part[1];
        sl[u] = (short)part[2];
    }
    h = __builtin_bit_cast( bf16x8, sh );
    m = __builtin_bit_cast( bf16x8, sm );
    l = __builtin_bit_cast( bf16x8, sl );
}

__device__ __forceinline__ bf16x8 load8( const unsigned short *p )
{
    return __builtin_bit_cast( bf16x8, *reinterpret_cast<const s16x8 *>( p ) );
}

// c += A B over one 32-deep K block, A from the three parts of the lane's
// 8 activations, B from the three weight parts at w[0], w[stride], w[2 stride].
__device__ __forceinline__ f32x4 mfma_split( const float *a8, const unsigned short *w,
                                             size_t stride, f32x4 c )
{
    bf16x8 ah, am, al;
    split8( a8, ah, am, al );
    const bf16x8 bh = load8( w ), bm = load8( w + stride ), bl = load8( w + 2 * stride );
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16( al, bh, c, 0, 0, 0 );
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16( ah, bl, c, 0, 0, 0 );
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16( am, bm, c, 0, 0, 0 );
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16( am, bh, c, 0, 0, 0 );
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16( ah, bm, c, 0, 0, 0 );
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16( ah, bh, c, 0, 0, 0 );
    return c;
}

// MFMA 16x16x32 operand layout (per lane l): A[row = l % 16][k = 8 (l / 16)
// + 0..7], B[k = 8 (l / 16) + 0..7][col = l % 16], D[row = 4 (l / 16) + 0..3]
// [col = l % 16].
__global__ __launch_bounds__( 256 ) void k_nnp_mlp( int elem, const int *__restrict__ list,
                                                   const int *__restrict__ count,
                                                   const float *__restrict__ G,
                                                   const NnpPlan *__restrict__ P,
                                                   float *__restrict__ dEdG,
                                                   float *__restrict__ Ei )
{
    __shared__ float sA[4][16][KPAD + 4]; // per wave: row-major 16 x K
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int t0 = ( blockIdx.x * 4 + wv ) * 16;
    const int cnt = count[elem];
    if ( t0 >= cnt )
        return;
    const NnpElemPlan &Q = P->el[elem];
    float( *A )[KPAD + 4] = sA[wv];
    const int row = lane & 15, kq = ( lane >> 4 ) * 4, k8 = ( lane >> 4 ) * 8;
    // scaled inputs, row-major in LDS
    for ( int k = lane; k < 16 * KPAD; k += 64 )
    {
        const int r = k / KPAD, c = k % KPAD;
        float v = 0.0f;
        if ( t0 + r < cnt && c < Q.nsf )
        {
            const int at = list[t0 + r];
            v = Q.g_mul[c] * G[(size_t)at * NSF_LD + c] + Q.g_add[c];
        }
        A[r][c] = v;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // layer 1: h1[16 x 32] = Gs[16 x 96] W1^T, B[k][col] = W1[col][k]
    f32x4 h1[2];
#pragma unroll
    for ( int nt = 0; nt < 2; ++nt )
    {
        f32x4 c = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for ( int kb = 0; kb < KPAD / 32; ++kb )
            c = mfma_split( &A[row][kb * 32 + k8], &Q.W1s[0][nt * 16 + row][kb * 32 + k8],
                            (size_t)HPAD * KPAD, c );
        h1[nt] = c;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // bias + activation; h1 to LDS (row-major 16 x 32)
#pragma unroll
    for ( int nt = 0; nt < 2; ++nt )
#pragma unroll
        for ( int u = 0; u < 4; ++u )
        {
            const int r = kq + u, col = nt * 16 + row;
            const float z = h1[nt][u] + ( col < Q.h1 ? Q.b1[col] : 0.f );
            const float h = col < Q.h1 ? act_f( z, Q.act1 ) : 0.f;
            h1[nt][u] = h;
            A[r][col] = h;
        }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // layer 2: h2[16 x 32] = h1 W2^T, B[k][col] = W2[col][k]
    f32x4 h2[2];
#pragma unroll
    for ( int nt = 0; nt < 2; ++nt )
        h2[nt] = mfma_split( &A[row][k8], &Q.W2s[0][nt * 16 + row][k8], (size_t)HPAD * HPAD,
                             f32x4{ 0.f, 0.f, 0.f, 0.f } );
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // output layer and the first backward step (VALU, FP32):
    // E = w3 . h2 + b3; delta2 = w3 * act'(z2)
    float epart[4] = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for ( int nt = 0; nt < 2; ++nt )
#pragma unroll
        for ( int u = 0; u < 4; ++u )
        {
            const int r = kq + u, col = nt * 16 + row;
            const float z = h2[nt][u] + ( col < Q.h2 ? Q.b2[col] : 0.f );
            const float h = col < Q.h2 ? act_f( z, Q.act2 ) : 0.f;
            const float w = col < Q.h2 ? Q.w3[col] : 0.f;
            epart[u] += h * w;
            A[r][col] = Q.act2 ? w * ( 1.f - h * h ) : w; // delta2, row-major
        }
    // sum the 16 columns held by the 16 lanes of each row group
#pragma unroll
    for ( int u = 0; u < 4; ++u )
    {
        float s = epart[u];
#pragma unroll
        for ( int o = 1; o < 16; o <<= 1 )
            s += __shfl_xor( s, o, 64 );
        epart[u] = s;
    }
    if ( row == 0 )
#pragma unroll
        for ( int u = 0; u < 4; ++u )
            if ( t0 + kq + u < cnt )
                Ei[list[t0 + kq + u]] = epart[u] + Q.b3;
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // delta1[16 x 32] = (delta2 W2) * act'(z1): B[k][col] = W2[k][col]
    f32x4 d1[2];
#pragma unroll
    for ( int nt = 0; nt < 2; ++nt )
        d1[nt] = mfma_split( &A[row][k8], &Q.W2Ts[0][nt * 16 + row][k8], (size_t)HPAD * HPAD,
                             f32x4{ 0.f, 0.f, 0.f, 0.f } );
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
#pragma unroll
    for ( int nt = 0; nt < 2; ++nt )
#pragma unroll
        for ( int u = 0; u < 4; ++u )
        {
            const int r = kq + u, col = nt * 16 + row;
            const float h = h1[nt][u];
            A[r][col] = col < Q.h1 ? ( Q.act1 ? d1[nt][u] * ( 1.f - h * h ) : d1[nt][u] ) : 0.f;
        }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // dE/dGs[16 x 96] = delta1 W1: B[k][col] = W1[k][col]
#pragma unroll
    for ( int nt = 0; nt < KPAD / 16; ++nt )
    {
        const f32x4 c = mfma_split( &A[row][k8], &Q.W1Ts[0][nt * 16 + row][k8],
                                    (size_t)KPAD * HPAD, f32x4{ 0.f, 0.f, 0.f, 0.f } );
        const int col = nt * 16 + row;
#pragma unroll
        for ( int u = 0; u < 4; ++u )
            if ( t0 + kq + u < cnt && col < Q.nsf )
                dEdG[(size_t)list[t0 + kq + u] * NSF_LD + col] = c[u] * Q.g_mul[col];
    }
}

// The same network in FP32 on the VALU, one atom per lane.
__global__ __launch_bounds__( 256 ) void k_nnp_mlp_f32( int n_local,
                                                       const double4 *__restrict__ x,
                                                       const float *__restrict__ G,
                                                       const NnpPlan *__restrict__ P,
                                                       float *__restrict__ dEdG,
                                                       float *__restrict__ Ei )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n_local )
        return;
    const NnpElemPlan &Q = P->el[(int)x[i].w];
    const float *Gi = G + (size_t)i * NSF_LD;
    float gs[KPAD];
    for ( int k = 0; k < Q.nsf; ++k )
        gs[k] = Q.g_mul[k] * Gi[k] + Q.g_add[k];
    float h1[HPAD], h2[HPAD];
    for ( int o = 0; o < Q.h1; ++o )
    {
        float z = Q.b1[o];
        for ( int k = 0; k < Q.nsf; ++k )
            z = fmaf( Q.W1[o][k], gs[k], z );
        h1[o] = act_f( z, Q.act1 );
    }
    float e = Q.b3;
    for ( int o = 0; o < Q.h2; ++o )
    {
        float z = Q.b2[o];
        for ( int k = 0; k < Q.h1; ++k )
            z = fmaf( Q.W2[o][k], h1[k], z );
        h2[o] = act_f( z, Q.act2 );
        e = fmaf( Q.w3[o], h2[o], e );
    }
    Ei[i] = e;
    float d2[HPAD], d1[HPAD];
    for ( int o = 0; o < Q.h2; ++o )
        d2[o] = Q.act2 ? Q.w3[o] * ( 1.f - h2[o] * h2[o] ) : Q.w3[o];
    for ( int k = 0; k < Q.h1; ++k )
    {
        float s = 0.f;
        for ( int o = 0; o < Q.h2; ++o )
            s = fmaf( Q.W2[o][k], d2[o], s );
        d1[k] = Q.act1 ? s * ( 1.f - h1[k] * h1[k] ) : s;
    }
    float *Di = dEdG + (size_t)i * NSF_LD;
    for ( int k = 0; k < Q.nsf; ++k )
    {
        float s = 0.f;
        for ( int o = 0; o < Q.h1; ++o )
            s = fmaf( Q.W1[o][k], d1[o], s );
        Di[k] = s * Q.g_mul[k];
    }
}

// ---- forces
template <int NAG, int NAP, bool RS0>
__global__ __launch_bounds__( NNP_TPB ) void k_nnp_force(
    int n_local, int fcap, const double4 *__restrict__ x,
    const int64_t *__restrict__ off, const int *__restrict__ idx,
    const NnpPlan *__restrict__ P, const float *__restrict__ dEdG,
    double *__restrict__ f, int *__restrict__ flag )
{
    __shared__ NbLds lds[NNP_TPB / 64];
    __shared__ float sF[NNP_TPB / 64][NN_MAX][3];
    __shared__ float sC[NNP_TPB / 64][NSF_LD];
    __shared__ float4 sCg[NNP_TPB / 64][NA_MAX / 4]; // c of angular (g, p) in grid order
    __shared__ unsigned tbuf[NNP_TPB / 64][TCAP];
    __shared__ float kst[NNP_TPB / 64][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * ( NNP_TPB / 64 ) + wv;
    if ( i >= n_local )
        return;
    NbLds &L = lds[wv];
    const float rc = P->rc, rc2 = P->rc2, kpi = P->pi_rc;
    const int n = gather_neighbours( i, x, off, idx, rc, rc2, kpi, L, flag );
    const int e = __builtin_amdgcn_readfirstlane( (int)x[i].w );
    const NnpElemPlan &Q = P->el[e];
    AngK K = load_angk( Q, kst[wv] );
    if constexpr ( NAG > 0 )
        K.nag = NAG;
    if constexpr ( NAP > 0 )
        K.nap = NAP;
    for ( int k = lane; k < NSF_LD; k += 64 )
        sC[wv][k] = k < Q.nsf ? dEdG[(size_t)i * NSF_LD + k] : 0.f;
    for ( int a = lane; a < n; a += 64 )
    {
        sF[wv][a][0] = 0.f;
        sF[wv][a][1] = 0.f;
        sF[wv][a][2] = 0.f;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    const float *c = sC[wv];
    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    // radial
    for ( int a = lane; a < n; a += 64 )
    {
        const float4 A = L.d[a], Ea = L.e[a];
        const float r = A.w, fc = Ea.y, dfc = Ea.z;
        const int ej = __builtin_bit_cast( int, Ea.w );
        float gsum = 0.f; // sum_s c_s dG_s/dr
#pragma unroll
        for ( int g = 0; g < RG_MAX; ++g )
        {
            if ( g < Q.nrg )
            {
                const float t = r - Q.r_rs[g];
                const float ex = __expf( -Q.r_eta[g] * t * t );
                const float d = ex * ( -2.f * Q.r_eta[g] * t * fc + dfc );
                float cw = 0.f;
#pragma unroll
                for ( int k = 0; k < RE_MAX; ++k )
                    if ( k < Q.nre && ( Q.r_el[k] < 0 || Q.r_el[k] == ej ) )
                        cw += c[Q.r_out[g][k]];
                gsum = fmaf( cw, d, gsum );
            }
        }
        const float s = gsum * Ea.x;
        const float vx = s * A.x, vy = s * A.y, vz = s * A.z;
        // dG/dx_j = g' u; F_j -= c g' u, F_i += c g' u
        sF[wv][a][0] -= vx;
        sF[wv][a][1] -= vy;
        sF[wv][a][2] -= vz;
        fix += vx;
        fiy += vy;
        fiz += vz;
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    // angular: coefficients in grid order in LDS (the generic grid reads
    // them per pair, which keeps 64 registers free for occupancy)
    {
        float *cg = reinterpret_cast<float *>( sCg[wv] );
        for ( int k = lane; k < NA_MAX; k += 64 )
        {
            const int g = k / AP_MAX, pp = k % AP_MAX;
            cg[k] = ( g < Q.nag && pp < Q.nap ) ? c[Q.a_out[g][pp]] : 0.f;
        }
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    const float4 *cg4 = sCg[wv];
    // with the grid fixed at compile time the coefficients are read once per
    // atom as wave-uniform values (scalar registers, the excess spilled to
    // lanes of vector registers) instead of from LDS per triplet: 4 % off
    // the kernel at 192,000 atoms
    float cgr[( NAG > 0 ? NAG : 1 ) * ( NAP > 0 ? NAP : 1 )];
    if constexpr ( NAG > 0 && NAP > 0 )
    {
        const float *cgf = reinterpret_cast<const float *>( sCg[wv] );
#pragma unroll
        for ( int g = 0; g < NAG; ++g )
#pragma unroll
            for ( int pp = 0; pp < NAP; ++pp )
                cgr[g * NAP + pp] = __builtin_bit_cast(
                    float, __builtin_amdgcn_readfirstlane(
                               __builtin_bit_cast( int, cgf[g * AP_MAX + pp] ) ) );
    }
    // angular: each lane takes a contiguous block of the compacted (a-major)
    // pair list, so the reaction on a stays in registers until a changes;
    // the one on q goes to LDS.
    int cur = -1;
    float ax = 0.f, ay = 0.f, az = 0.f; // reaction on row atom cur
    // gradients of sum_s c_s T_s of the triplet (i, a, q) with respect to
    // A = x_a - x_i (ga) and B = x_q - x_i (gb): no side effects, so two of
    // them run interleaved
    auto grad = [&]( int a, int q, float ( &ga )[3], float ( &gb )[3] ) {
        const float4 A = L.d[a], Ea = L.e[a];
        const float4 B = L.d[q], Eb = L.e[q];
        const float cx = B.x - A.x, cy = B.y - A.y, cz = B.z - A.z;
        const float r1 = A.w, r2 = B.w;
        const float r3 = __builtin_amdgcn_sqrtf( cx * cx + cy * cy + cz * cz );
        const float ia = Ea.x, ib = Eb.x, ic = __builtin_amdgcn_rcpf( r3 );
        const float f1 = Ea.y, f2 = Eb.y, f3 = fcut( r3, rc, kpi );
        const float d1 = Ea.z, d2 = Eb.z, d3 = dfcut( r3, rc, kpi );
        const float F = f1 * f2 * f3;
        const float inv12 = ia * ib;
        const float cs = ( A.x * B.x + A.y * B.y + A.z * B.z ) * inv12;
        float Pw[AP_MAX], Dw[AP_MAX];
        ang_powers<true>( K, cs, Pw, Dw );
        float Gs = 0.f, As = 0.f, B1 = 0.f, B2 = 0.f, B3 = 0.f;
        // sum_p c_gp P_p and sum_p c_gp D_p of one (eta, rs)
        auto cdots = [&]( int g, float &sp, float &sd )
        {
            sp = 0.f;
            sd = 0.f;
            if constexpr ( NAG > 0 && NAP > 0 )
            {
#pragma unroll
                for ( int pp = 0; pp < NAP; ++pp )
                {
                    sp = fmaf( cgr[g * NAP + pp], Pw[pp], sp );
                    sd = fmaf( cgr[g * NAP + pp], Dw[pp], sd );
                }
            }
            else
            {
#pragma unroll
                for ( int q4 = 0; q4 < AP_MAX / 4; ++q4 )
                {
                    const float4 cc = cg4[g * ( AP_MAX / 4 ) + q4];
                    const float cv[4] = { cc.x, cc.y, cc.z, cc.w };
#pragma unroll
                    for ( int u = 0; u < 4; ++u )
                    {
                        sp = fmaf( cv[u], Pw[q4 * 4 + u], sp );
                        sd = fmaf( cv[u], Dw[q4 * 4 + u], sd );
                    }
                }
            }
        };
        if constexpr ( RS0 )
        {
            // one Gaussian argument, and the distance gradients share one
            // factor, sum_g -2 eta_g w_g
            const float s2 = r1 * r1 + r2 * r2 + r3 * r3;
            float M2 = 0.f;
#pragma unroll
            for ( int g = 0; g < AG_MAX; ++g )
            {
                if ( g < K.nag )
                {
                    const float Eg = __expf( -K.eta[g] * s2 );
                    float sp, sd;
                    cdots( g, sp, sd );
                    const float w = sp * Eg; // sum_p c P E
                    Gs += w;
                    As = fmaf( sd, Eg, As );
                    M2 = fmaf( -2.f * K.eta[g], w, M2 );
                }
            }
            B1 = M2 * r1;
            B2 = M2 * r2;
            B3 = M2 * r3;
        }
        else
        {
#pragma unroll
            for ( int g = 0; g < AG_MAX; ++g )
            {
                if ( g < K.nag )
                {
                    const float rs = K.rs[g];
                    const float t1 = r1 - rs, t2 = r2 - rs, t3 = r3 - rs;
                    const float Eg = __expf( -K.eta[g] * ( t1 * t1 + t2 * t2 + t3 * t3 ) );
                    float sp, sd;
                    cdots( g, sp, sd );
                    const float w = sp * Eg; // sum_p c P E
                    Gs += w;
                    As = fmaf( sd, Eg, As );
                    const float m2 = -2.f * K.eta[g] * w;
                    B1 = fmaf( m2, t1, B1 );
                    B2 = fmaf( m2, t2, B2 );
                    B3 = fmaf( m2, t3, B3 );
                }
            }
        }
        // (T = pref P E F, r3 = |B - A|)
        const float dFa1 = d1 * f2 * f3, dF3 = f1 * f2 * d3, dFb2 = f1 * d2 * f3;
        const float Av[3] = { A.x, A.y, A.z }, Bv[3] = { B.x, B.y, B.z },
                    Cv[3] = { cx, cy, cz };
#pragma unroll
        for ( int k = 0; k < 3; ++k )
        {
            const float u1 = Av[k] * ia, u2 = Bv[k] * ib, u3 = Cv[k] * ic;
            const float dca = Bv[k] * inv12 - cs * Av[k] * ia * ia;
            const float dcb = Av[k] * inv12 - cs * Bv[k] * ib * ib;
            ga[k] = F * ( As * dca + B1 * u1 - B3 * u3 ) + Gs * ( dFa1 * u1 - dF3 * u3 );
            gb[k] = F * ( As * dcb + B2 * u2 + B3 * u3 ) + Gs * ( dFb2 * u2 + dF3 * u3 );
        }
    };
    // the reactions of one triplet: on a in registers while the row atom
    // stays the same (the stream is a-major), on q through LDS
    auto apply = [&]( int a, int q, const float ( &ga )[3], const float ( &gb )[3] ) {
        if ( a != cur )
        {
            if ( cur >= 0 )
            {
                atomicAdd( &sF[wv][cur][0], ax );
                atomicAdd( &sF[wv][cur][1], ay );
                atomicAdd( &sF[wv][cur][2], az );
            }
            cur = a;
            ax = ay = az = 0.f;
        }
        ax -= ga[0];
        ay -= ga[1];
        az -= ga[2];
        atomicAdd( &sF[wv][q][0], -gb[0] );
        atomicAdd( &sF[wv][q][1], -gb[1] );
        atomicAdd( &sF[wv][q][2], -gb[2] );
        fix += ga[0] + gb[0];
        fiy += ga[1] + gb[1];
        fiz += ga[2] + gb[2];
    };
    if constexpr ( NAG > 0 && NAP > 0 )
        for_triplets<true>( L, n, rc2, tbuf[wv],
                            [&]( int a0, int q0, int a1, int q1, bool ok1 ) {
                                float ga0[3], gb0[3], ga1[3], gb1[3];
                                grad( a0, q0, ga0, gb0 );
                                grad( a1, q1, ga1, gb1 );
                                apply( a0, q0, ga0, gb0 );
                                if ( ok1 )
                                    apply( a1, q1, ga1, gb1 );
                            } );
    else
        // the generic grid's registers leave no room for a second chain
        for_triplets( L, n, rc2, tbuf[wv], [&]( int a, int q ) {
            float ga[3], gb[3];
            grad( a, q, ga, gb );
            apply( a, q, ga, gb );
        } );
    if ( cur >= 0 )
    {
        atomicAdd( &sF[wv][cur][0], ax );
        atomicAdd( &sF[wv][cur][1], ay );
        atomicAdd( &sF[wv][cur][2], az );
    }
    __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
    {
        fix += __shfl_xor( fix, o, 64 );
        fiy += __shfl_xor( fiy, o, 64 );
        fiz += __shfl_xor( fiz, o, 64 );
    }
    if ( lane == 0 )
    {
        unsafeAtomicAdd( &f[i], (double)fix );
        unsafeAtomicAdd( &f[fcap + i], (double)fiy );
        unsafeAtomicAdd( &f[2 * fcap + i], (double)fiz );
    }
    for ( int a = lane; a < n; a += 64 )
    {
        const int j = L.j[a];
        unsafeAtomicAdd( &f[j], (double)sF[wv][a][0] );
        unsafeAtomicAdd( &f[fcap + j], (double)sF[wv][a][1] );
        unsafeAtomicAdd( &f[2 * fcap + j], (double)sF[wv][a][2] );
    }
}

// Atoms grouped by element for the per-element network GEMMs; one atomic
// per element and wave (the order inside a group does not matter).
__global__ void k_nnp_group( int n_local, const double4 *__restrict__ x,
                             int cap, int *__restrict__ list, int *__restrict__ count )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int e = i < n_local ? (int)x[i].w : -1;
    for ( int el = 0; el < NNP_MAXEL; ++el )
    {
        const unsigned long long m = __ballot( e == el );
        if ( !m )
            continue;
        const int leader = __ffsll( (long long)m ) - 1;
        int base = 0;
        if ( lane == leader )
            base = atomicAdd( &count[el], (int)__popcll( m ) );
        base = __shfl( base, leader, 64 );
        if ( e == el )
        {
            const int r = (int)__builtin_amdgcn_mbcnt_hi(
                (unsigned)( m >> 32 ), __builtin_amdgcn_mbcnt_lo( (unsigned)m, 0u ) );
            list[(size_t)el * cap + base + r] = i;
        }
    }
}

__global__ void k_nnp_esum( int n, const float *__restrict__ Ei,
                            double *__restrict__ part )
{
    __shared__ double s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    s[threadIdx.x] = i < n ? (double)Ei[i] : 0.0;
    __syncthreads();
    for ( int o = 128; o > 0; o >>= 1 )
    {
        if ( (int)threadIdx.x < o )
            s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        part[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------- host
static cmd_status nnp_prepare( cmd_nnp *M, int n_local, hipStream_t st )
{
    if ( !M->plan_ok )
    {
        CMD_TRY( build_plan( M->model, M->plan ) );
        CMD_TRY( M->dplan.reserve( 1 ) );
        CMD_HIP( hipMemcpyAsync( M->dplan.p, &M->plan, sizeof( NnpPlan ),
                                 hipMemcpyHostToDevice, st ) );
        M->plan_ok = true;
    }
    const int n = std::max( n_local, 1 );
    if ( n > M->cap )
    {
        M->cap = n;
        CMD_TRY( M->G.reserve( (size_t)n * NSF_LD ) );
        CMD_TRY( M->dEdG.reserve( (size_t)n * NSF_LD ) );
        CMD_TRY( M->Ei.reserve( n ) );
        CMD_TRY( M->elist.reserve( (size_t)n * NNP_MAXEL ) );
    }
    CMD_TRY( M->ecount.reserve( NNP_MAXEL ) );
    CMD_TRY( M->flag.reserve( 1 ) );
    return CMD_OK;
}

cmd_status nnp_compute( cmd_nnp *M, cmd_system *S, cmd_neigh *N )
{
    const hipStream_t st = S->stream;
    const int n = S->n_local;
    CMD_TRY( nnp_prepare( M, n, st ) );
    CMD_REQUIRE( S->ntypes <= M->plan.nelem, CMD_ERR_ARG,
                 "the system has %d atom types but the model %d elements",
                 S->ntypes, M->plan.nelem );
    CMD_REQUIRE( !N->half, CMD_ERR_ARG,
                 "the NNP force needs a full neighbour list (the reference "
                 "supports only full lists, force_nnp_cabana_neigh_impl.h)" );
    // A list shorter than the symmetry-function cutoff truncates the
    // functions to the listed neighbours, as in the reference (in.nnp lists
    // 3.9 + skin against rc 7.37 in input/nnp/input.nn).
    CMD_TRY( neigh_csr( N ) );
    const int nt = S->n_local + S->n_ghost;
    for ( int c = 0; c < 3; ++c )
        CMD_TRY( fill_f64( S->f.p + (size_t)c * S->cap, 0.0, nt, st ) );
    CMD_TRY( fill_i32( M->flag.p, 0, 1, st ) );
    if ( n == 0 )
        return CMD_OK;
    CMD_TRY( fill_i32( reinterpret_cast<int *>( M->G.p ), 0, (size_t)n * NSF_LD, st ) );
    CMD_TRY( M->terms.reserve( std::max( n, 1 ) ) );
    M->n_last = n;
    const int nbw = ( n + NNP_TPB / 64 - 1 ) / ( NNP_TPB / 64 );
    // the angular grid of every element, for the kernels compiled for it:
    // its size (5 x 8) and whether all its rs are 0
    bool g58 = M->plan.nelem > 0, rs0 = true;
    for ( int e = 0; e < M->plan.nelem; ++e )
    {
        const NnpElemPlan &Q = M->plan.el[e];
        g58 = g58 && Q.nag == 5 && Q.nap == 8;
        for ( int g = 0; g < Q.nag; ++g )
            rs0 = rs0 && Q.a_rs[g] == 0.f;
    }
#define CMD_SF( A, P, R )                                                                 \
    k_nnp_sf<A, P, R><<<nbw, NNP_TPB, 0, st>>>( n, S->x.p, N->offsets.p, N->idx.p, M->dplan.p, \
                                                M->G.p, M->flag.p, M->terms.p )
    if ( g58 )
    {
        if ( rs0 )
            CMD_SF( 5, 8, true );
        else
            CMD_SF( 5, 8, false );
    }
    else if ( rs0 )
        CMD_SF( 0, 0, true );
    else
        CMD_SF( 0, 0, false );
#undef CMD_SF
    CMD_LAUNCH_CHECK();
    if ( M->prec == CMD_PREC_MIXED )
    {
        CMD_TRY( fill_i32( M->ecount.p, 0, NNP_MAXEL, st ) );
        k_nnp_group<<<( n + 255 ) / 256, 256, 0, st>>>( n, S->x.p, M->cap,
                                                       M->elist.p, M->ecount.p );
        CMD_LAUNCH_CHECK();
        for ( int e = 0; e < M->plan.nelem; ++e )
        {
            k_nnp_mlp<<<( n + 63 ) / 64, 256, 0, st>>>(
                e, M->elist.p + (size_t)e * M->cap, M->ecount.p, M->G.p,
                M->dplan.p, M->dEdG.p, M->Ei.p );
            CMD_LAUNCH_CHECK();
        }
    }
    else
    {
        k_nnp_mlp_f32<<<( n + 255 ) / 256, 256, 0, st>>>( n, S->x.p, M->G.p,
                                                         M->dplan.p, M->dEdG.p,
                                                         M->Ei.p );
        CMD_LAUNCH_CHECK();
    }
#define CMD_FORCE( A, P, R )                                                               \
    k_nnp_force<A, P, R><<<nbw, NNP_TPB, 0, st>>>( n, S->cap, S->x.p, N->offsets.p, N->idx.p, \
                                                   M->dplan.p, M->dEdG.p, S->f.p, M->flag.p )
    if ( g58 )
    {
        if ( rs0 )
            CMD_FORCE( 5, 8, true );
        else
            CMD_FORCE( 5, 8, false );
    }
    else if ( rs0 )
        CMD_FORCE( 0, 0, true );
    else
        CMD_FORCE( 0, 0, false );
#undef CMD_FORCE
    CMD_LAUNCH_CHECK();
    int fl = 0;
    CMD_HIP( hipMemcpyAsync( &fl, M->flag.p, sizeof( int ),
                             hipMemcpyDeviceToHost, st ) );
    CMD_HIP( hipStreamSynchronize( st ) );
    CMD_REQUIRE( fl == 0, CMD_ERR_CAPACITY,
                 "an atom has more than %d neighbours inside the NNP cutoff",
                 NN_MAX );
    return CMD_OK;
}

// Sum of the atomic network outputs (FP64 reduction).
cmd_status nnp_energy_sum( cmd_nnp *M, cmd_system *S, double *sum )
{
    const int n = S->n_local;
    const int nb = ( n + 255 ) / 256;
    CMD_TRY( M->part.reserve( std::max( nb, 1 ) ) );
    if ( n > 0 )
    {
        k_nnp_esum<<<nb, 256, 0, S->stream>>>( n, M->Ei.p, M->part.p );
        CMD_LAUNCH_CHECK();
    }
    std::vector<double> h( std::max( nb, 1 ), 0.0 );
    if ( n > 0 )
        CMD_HIP( hipMemcpyAsync( h.data(), M->part.p, sizeof( double ) * nb,
                                 hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    double s = 0.0;
    for ( int k = 0; k < nb; ++k )
        s += h[k];
    *sum = s;
    return CMD_OK;
}

// ---------------------------------------------------------------- n2p2 files
static std::vector<std::string> words_of( const std::string &line )
{
    std::vector<std::string> w;
    std::istringstream is( line.substr( 0, line.find( '#' ) ) );
    std::string t;
    while ( is >> t )
        w.push_back( t );
    return w;
}

// n2p2 orders an element's functions by type, then cutoff, then the
// parameters (eta, rs, zeta, lambda) and neighbour elements; scaling.data
// rows and the network inputs follow that order.
static bool sf_less( const cmd_nnp_model::SF &a, const cmd_nnp_model::SF &b )
{
    if ( a.type != b.type )
        return a.type < b.type;
    if ( a.rc != b.rc )
        return a.rc < b.rc;
    if ( a.eta != b.eta )
        return a.eta < b.eta;
    if ( a.rs != b.rs )
        return a.rs < b.rs;
    if ( a.zeta != b.zeta )
        return a.zeta < b.zeta;
    if ( a.lambda != b.lambda )
        return a.lambda < b.lambda;
    if ( a.e1 != b.e1 )
        return a.e1 < b.e1;
    return a.e2 < b.e2;
}

static cmd_status load_n2p2( const char *dir, int nelem, const char *const *names,
                             const int *znum, cmd_nnp_model &m )
{
    m = cmd_nnp_model{};
    m.nelem = nelem;
    m.el.resize( nelem );
    m.scale = false;
    for ( int e = 0; e < nelem; ++e )
    {
        m.el[e].name = names[e];
        m.el[e].z = znum[e];
    }
    auto elem_of = [&]( const std::string &s ) -> int
    {
        for ( int e = 0; e < nelem; ++e )
            if ( m.el[e].name == s )
                return e;
        return -1;
    };
    const std::string d( dir );
    {
        std::ifstream in( d + "/input.nn" );
        CMD_REQUIRE( in.good(), CMD_ERR_IO, "cannot read %s/input.nn", dir );
        std::string line;
        std::string act;
        while ( std::getline( in, line ) )
        {
            const auto w = words_of( line );
            if ( w.empty() )
                continue;
            try
            {
                if ( w[0] == "global_nodes_short" )
                {
                    m.hidden.clear();
                    for ( size_t k = 1; k < w.size(); ++k )
                        m.hidden.push_back( std::stoi( w[k] ) );
                }
                else if ( w[0] == "global_activation_short" )
                {
                    act.clear();
                    for ( size_t k = 1; k < w.size(); ++k )
                        act += w[k];
                }
                else if ( w[0] == "scale_symmetry_functions" )
                    m.scale = true;
                else if ( w[0] == "scale_min_short" )
                    m.smin = std::stod( w.at( 1 ) );
                else if ( w[0] == "scale_max_short" )
                    m.smax = std::stod( w.at( 1 ) );
                else if ( w[0] == "cutoff_type" )
                    CMD_REQUIRE( std::stoi( w.at( 1 ) ) == 1, CMD_ERR_IO,
                                 "only cutoff_type 1 (cosine) is supported" );
                else if ( w[0] == "atom_energy" )
                {
                    const int e = elem_of( w.at( 1 ) );
                    if ( e >= 0 )
                        m.el[e].atom_energy = std::stod( w.at( 2 ) );
                }
                else if ( w[0] == "symfunction_short" )
                {
                    const int e = elem_of( w.at( 1 ) );
                    CMD_REQUIRE( e >= 0, CMD_ERR_IO, "unknown element %s",
                                 w.at( 1 ).c_str() );
                    cmd_nnp_model::SF s{};
                    s.type = std::stoi( w.at( 2 ) );
                    if ( s.type == 2 )
                    {
                        s.e1 = elem_of( w.at( 3 ) );
                        s.e2 = -1;
                        s.eta = std::stod( w.at( 4 ) );
                        s.rs = std::stod( w.at( 5 ) );
                        s.rc = std::stod( w.at( 6 ) );
                    }
                    else if ( s.type == 3 )
                    {
                        s.e1 = elem_of( w.at( 3 ) );
                        s.e2 = elem_of( w.at( 4 ) );
                        s.eta = std::stod( w.at( 5 ) );
                        s.lambda = std::stod( w.at( 6 ) );
                        s.zeta = std::stod( w.at( 7 ) );
                        s.rc = std::stod( w.at( 8 ) );
                        s.rs = w.size() > 9 ? std::stod( w[9] ) : 0.0;
                    }
                    else
                    {
                        set_error( "symfunction_short type %d is not supported",
                                   s.type );
                        return CMD_ERR_IO;
                    }
                    m.el[e].sf.push_back( s );
                }
            }
            catch ( const std::exception & )
            {
                set_error( "malformed line in input.nn: %s", line.c_str() );
                return CMD_ERR_IO;
            }
        }
        m.act = act;
        m.nlayers = (int)m.hidden.size() + 1;
    }
    for ( auto &E : m.el )
    {
        std::stable_sort( E.sf.begin(), E.sf.end(), sf_less );
        E.gmin.assign( E.sf.size(), 0.0 );
        E.gmax.assign( E.sf.size(), 1.0 );
    }
    if ( m.scale )
    {
        std::ifstream in( d + "/scaling.data" );
        CMD_REQUIRE( in.good(), CMD_ERR_IO, "cannot read %s/scaling.data", dir );
        std::string line;
        while ( std::getline( in, line ) )
        {
            const auto w = words_of( line );
            if ( w.size() < 4 )
                continue;
            const int e = std::stoi( w[0] ) - 1, k = std::stoi( w[1] ) - 1;
            CMD_REQUIRE( e >= 0 && e < nelem && k >= 0 &&
                             k < (int)m.el[e].sf.size(),
                         CMD_ERR_IO, "scaling.data row %d %d out of range", e + 1,
                         k + 1 );
            m.el[e].gmin[k] = std::stod( w[2] );
            m.el[e].gmax[k] = std::stod( w[3] );
        }
    }
    for ( int e = 0; e < nelem; ++e )
    {
        auto &E = m.el[e];
        std::vector<int> sizes;
        sizes.push_back( (int)E.sf.size() );
        for ( int h : m.hidden )
            sizes.push_back( h );
        sizes.push_back( 1 );
        const int L = (int)sizes.size() - 1;
        E.W.assign( L, {} );
        E.b.assign( L, {} );
        for ( int l = 0; l < L; ++l )
        {
            E.W[l].assign( (size_t)sizes[l + 1] * sizes[l], 0.0 );
            E.b[l].assign( sizes[l + 1], 0.0 );
        }
        char fn[64];
        std::snprintf( fn, sizeof( fn ), "/weights.%03d.data", E.z );
        std::ifstream in( d + fn );
        CMD_REQUIRE( in.good(), CMD_ERR_IO, "cannot read %s%s", dir, fn );
        std::string line;
        while ( std::getline( in, line ) )
        {
            const auto w = words_of( line );
            if ( w.size() < 5 )
                continue;
            try
            {
                const double v = std::stod( w[0] );
                if ( w[1] == "a" && w.size() >= 7 )
                {
                    const int lf = std::stoi( w[3] ), nf = std::stoi( w[4] ),
                              nt = std::stoi( w[6] );
                    CMD_REQUIRE( lf >= 0 && lf < L && nf >= 1 && nf <= sizes[lf] &&
                                     nt >= 1 && nt <= sizes[lf + 1],
                                 CMD_ERR_IO, "weight index out of range" );
                    E.W[lf][(size_t)( nt - 1 ) * sizes[lf] + ( nf - 1 )] = v;
                }
                else if ( w[1] == "b" )
                {
                    const int l = std::stoi( w[3] ), nd = std::stoi( w[4] );
                    CMD_REQUIRE( l >= 1 && l <= L && nd >= 1 && nd <= sizes[l],
                                 CMD_ERR_IO, "bias index out of range" );
                    E.b[l - 1][nd - 1] = v;
                }
            }
            catch ( const std::exception & )
            {
                set_error( "malformed line in %s: %s", fn + 1, line.c_str() );
                return CMD_ERR_IO;
            }
        }
    }
    return CMD_OK;
}

// ---------------------------------------------------------------- synthetic
// splitmix64 stream for seeded weights (normal deviates by Box-Muller).
struct Mix
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = ( s += 0x9e3779b97f4a7c15ull );
        z = ( z ^ ( z >> 30 ) ) * 0xbf58476d1ce4e5b9ull;
        z = ( z ^ ( z >> 27 ) ) * 0x94d049bb133111ebull;
        return z ^ ( z >> 31 );
    }
    double uniform() { return ( next() >> 11 ) * ( 1.0 / 9007199254740992.0 ); }
    double normal()
    {
        const double u1 = std::max( uniform(), 1e-300 ), u2 = uniform();
        return std::sqrt( -2.0 * std::log( u1 ) ) * std::cos( 6.283185307179586 * u2 );
    }
};

} // namespace cmd

using namespace cmd;

extern "C" {

cmd_status cmd_nnp_model_create( int nelem, int n_hidden, const int *hidden,
                                 const char *activations, cmd_nnp_model **out )
{
    CMD_REQUIRE( out && nelem >= 1 && nelem <= NNP_MAXEL && n_hidden >= 0 &&
                     activations && (int)std::strlen( activations ) == n_hidden + 1,
                 CMD_ERR_ARG, "bad NNP model description" );
    cmd_nnp_model *m = new cmd_nnp_model;
    m->nelem = nelem;
    m->el.resize( nelem );
    m->hidden.assign( hidden, hidden + n_hidden );
    m->nlayers = n_hidden + 1;
    m->act = activations;
    *out = m;
    return CMD_OK;
}

cmd_status cmd_nnp_model_destroy( cmd_nnp_model *m )
{
    delete m;
    return CMD_OK;
}

cmd_status cmd_nnp_model_set_scaling( cmd_nnp_model *m, int scale, double smin,
                                      double smax )
{
    CMD_REQUIRE( m, CMD_ERR_ARG, "model is NULL" );
    m->scale = scale != 0;
    m->smin = smin;
    m->smax = smax;
    return CMD_OK;
}

cmd_status cmd_nnp_model_add_sf( cmd_nnp_model *m, int elem, int type, int e1,
                                 int e2, double eta, double rs, double lambda,
                                 double zeta, double rc, double gmin, double gmax )
{
    CMD_REQUIRE( m && elem >= 0 && elem < m->nelem, CMD_ERR_ARG, "bad element" );
    CMD_REQUIRE( type == 2 || type == 3, CMD_ERR_ARG,
                 "symmetry function type %d is not supported", type );
    m->el[elem].sf.push_back( cmd_nnp_model::SF{ type, e1, e2, eta, rs, lambda, zeta, rc } );
    m->el[elem].gmin.push_back( gmin );
    m->el[elem].gmax.push_back( gmax );
    return CMD_OK;
}

cmd_status cmd_nnp_model_set_layer( cmd_nnp_model *m, int elem, int layer,
                                    int n_out, int n_in, const double *W,
                                    const double *b )
{
    CMD_REQUIRE( m && elem >= 0 && elem < m->nelem && layer >= 0 &&
                     layer < m->nlayers && W && b,
                 CMD_ERR_ARG, "bad layer" );
    auto &E = m->el[elem];
    if ( (int)E.W.size() < m->nlayers )
    {
        E.W.resize( m->nlayers );
        E.b.resize( m->nlayers );
    }
    E.W[layer].assign( W, W + (size_t)n_out * n_in );
    E.b[layer].assign( b, b + n_out );
    return CMD_OK;
}

cmd_status cmd_nnp_model_set_atom_energy( cmd_nnp_model *m, int elem, double e )
{
    CMD_REQUIRE( m && elem >= 0 && elem < m->nelem, CMD_ERR_ARG, "bad element" );
    m->el[elem].atom_energy = e;
    return CMD_OK;
}

cmd_status cmd_nnp_model_load_n2p2( const char *dir, int nelem,
                                    const char *const *names,
                                    const int *atomic_numbers,
                                    cmd_nnp_model **out )
{
    CMD_REQUIRE( dir && names && atomic_numbers && out && nelem >= 1 &&
                     nelem <= NNP_MAXEL,
                 CMD_ERR_ARG, "bad argument" );
    cmd_nnp_model *m = new cmd_nnp_model;
    const cmd_status st = load_n2p2( dir, nelem, names, atomic_numbers, *m );
    if ( st != CMD_OK )
    {
        delete m;
        return st;
    }
    *out = m;
    return CMD_OK;
}

// The benchmark model: 2 elements; per element 15 radial functions (eta
// log-spaced in [0.01, 2], rs = 0) for each neighbour element and 40
// angular ones (5 eta log-spaced in [0.001, 0.05] x lambda +-1 x zeta
// {1, 2, 4, 16}, any pair); cosine cutoff rc; network 70 -> 25 -> 25 -> 1
// (tanh, tanh, linear) with N(0, 1)/sqrt(fan_in) weights and zero biases
// from `seed`; scaling ranges [0, 1] until fitted to a frame.
cmd_status cmd_nnp_model_synthetic( uint64_t seed, double rc, cmd_nnp_model **out )
{
    CMD_REQUIRE( out && rc > 0, CMD_ERR_ARG, "bad argument" );
    cmd_nnp_model *m = new cmd_nnp_model;
    m->nelem = 2;
    m->el.resize( 2 );
    m->hidden = { 25, 25 };
    m->nlayers = 3;
    m->act = "ttl";
    m->scale = true;
    Mix rng{ seed };
    for ( int e = 0; e < 2; ++e )
    {
        auto &E = m->el[e];
        for ( int k = 0; k < 15; ++k )
        {
            const double eta = 0.01 * std::pow( 200.0, k / 14.0 );
            for ( int ne = 0; ne < 2; ++ne )
                E.sf.push_back( { 2, ne, -1, eta, 0.0, 0.0, 0.0, rc } );
        }
        const double zetas[4] = { 1.0, 2.0, 4.0, 16.0 };
        for ( int k = 0; k < 5; ++k )
        {
            const double eta = 0.001 * std::pow( 50.0, k / 4.0 );
            for ( double z : zetas )
                for ( double lam : { -1.0, 1.0 } )
                    E.sf.push_back( { 3, -1, -1, eta, 0.0, lam, z, rc } );
        }
        E.gmin.assign( E.sf.size(), 0.0 );
        E.gmax.assign( E.sf.size(), 1.0 );
        const int sizes[4] = { (int)E.sf.size(), 25, 25, 1 };
        E.W.resize( 3 );
        E.b.resize( 3 );
        for ( int l = 0; l < 3; ++l )
        {
            E.W[l].resize( (size_t)sizes[l + 1] * sizes[l] );
            for ( auto &w : E.W[l] )
                w = rng.normal() / std::sqrt( (double)sizes[l] );
            E.b[l].assign( sizes[l + 1], 0.0 );
        }
    }
    *out = m;
    return CMD_OK;
}

cmd_status cmd_nnp_model_info( const cmd_nnp_model *m, int elem, int *nsf,
                               int *n_hidden, int *hidden, double *atom_energy )
{
    CMD_REQUIRE( m && elem >= 0 && elem < m->nelem, CMD_ERR_ARG, "bad element" );
    if ( nsf )
        *nsf = (int)m->el[elem].sf.size();
    if ( n_hidden )
        *n_hidden = (int)m->hidden.size();
    if ( hidden )
        for ( size_t k = 0; k < m->hidden.size(); ++k )
            hidden[k] = m->hidden[k];
    if ( atom_energy )
        *atom_energy = m->el[elem].atom_energy;
    return CMD_OK;
}

// params: type, e1, e2, eta, rs, lambda, zeta, rc, gmin, gmax
cmd_status cmd_nnp_model_get_sf( const cmd_nnp_model *m, int elem, int k,
                                 double params[10] )
{
    CMD_REQUIRE( m && elem >= 0 && elem < m->nelem && k >= 0 &&
                     k < (int)m->el[elem].sf.size() && params,
                 CMD_ERR_ARG, "bad argument" );
    const auto &s = m->el[elem].sf[k];
    const double v[10] = { (double)s.type, (double)s.e1, (double)s.e2, s.eta, s.rs,
                           s.lambda, s.zeta, s.rc, m->el[elem].gmin[k],
                           m->el[elem].gmax[k] };
    std::memcpy( params, v, sizeof( v ) );
    return CMD_OK;
}

cmd_status cmd_nnp_model_get_layer( const cmd_nnp_model *m, int elem, int layer,
                                    double *W, double *b )
{
    CMD_REQUIRE( m && elem >= 0 && elem < m->nelem && layer >= 0 &&
                     layer < (int)m->el[elem].W.size(),
                 CMD_ERR_ARG, "bad argument" );
    if ( W )
        std::memcpy( W, m->el[elem].W[layer].data(),
                     sizeof( double ) * m->el[elem].W[layer].size() );
    if ( b )
        std::memcpy( b, m->el[elem].b[layer].data(),
                     sizeof( double ) * m->el[elem].b[layer].size() );
    return CMD_OK;
}

cmd_status cmd_force_nnp_create( const cmd_nnp_model *m, cmd_force **out )
{
    CMD_REQUIRE( m && out, CMD_ERR_ARG, "NULL argument" );
    cmd_force *f = new cmd_force;
    f->kind = 2;
    f->ntypes = m->nelem;
    f->nnp = new cmd_nnp;
    f->nnp->model = *m;
    std::vector<NnpPlan> probe( 1 ); // a plan holds the split weights: keep it off the stack
    const cmd_status st = build_plan( f->nnp->model, probe[0] );
    if ( st != CMD_OK )
    {
        delete f->nnp;
        delete f;
        return st;
    }
    f->prec = CMD_PREC_MIXED;
    *out = f;
    return CMD_OK;
}

// Scaling ranges fitted to the symmetry functions of the current frame
// (the benchmark model's "min-max scaling from the initial frame").
cmd_status cmd_force_nnp_fit_scaling( cmd_force *f, cmd_system *s, cmd_neigh *n )
{
    CMD_REQUIRE( f && f->kind == 2 && s && n, CMD_ERR_ARG, "bad argument" );
    cmd_nnp *M = f->nnp;
    const int nl = s->n_local;
    // raw symmetry functions with unit scaling
    cmd_nnp_model saved = M->model;
    for ( auto &E : M->model.el )
    {
        E.gmin.assign( E.sf.size(), 0.0 );
        E.gmax.assign( E.sf.size(), 1.0 );
    }
    M->model.smin = 0.0;
    M->model.smax = 1.0;
    M->plan_ok = false;
    cmd_status st = nnp_compute( M, s, n );
    std::vector<float> G( (size_t)std::max( nl, 1 ) * NSF_LD );
    std::vector<double4> x( std::max( nl, 1 ) );
    if ( st == CMD_OK && nl > 0 )
    {
        if ( hipMemcpy( G.data(), M->G.p, sizeof( float ) * G.size(),
                        hipMemcpyDeviceToHost ) != hipSuccess ||
             hipMemcpy( x.data(), s->x.p, sizeof( double4 ) * nl,
                        hipMemcpyDeviceToHost ) != hipSuccess )
        {
            set_error( "copy of symmetry functions failed" );
            st = CMD_ERR_HIP;
        }
    }
    M->model = saved;
    if ( st != CMD_OK )
        return st;
    for ( int e = 0; e < M->model.nelem; ++e )
    {
        auto &E = M->model.el[e];
        const int ns = (int)E.sf.size();
        std::vector<double> lo( ns, 1e300 ), hi( ns, -1e300 );
        for ( int i = 0; i < nl; ++i )
            if ( (int)x[i].w == e )
                for ( int k = 0; k < ns; ++k )
                {
                    lo[k] = std::min( lo[k], (double)G[(size_t)i * NSF_LD + k] );
                    hi[k] = std::max( hi[k], (double)G[(size_t)i * NSF_LD + k] );
                }
        for ( int k = 0; k < ns; ++k )
            if ( lo[k] <= hi[k] )
            {
                E.gmin[k] = lo[k];
                E.gmax[k] = hi[k];
            }
    }
    M->model.scale = true;
    M->plan_ok = false;
    return CMD_OK;
}

cmd_status cmd_nnp_lattice_create( int nx, int ny, int nz, double s,
                                   double jitter, double frac1, uint64_t seed,
                                   int *n, double *x, int *type, int *id )
{
    CMD_REQUIRE( nx > 0 && ny > 0 && nz > 0 && s > 0 && n && jitter >= 0 &&
                     jitter < 0.5 && frac1 >= 0 && frac1 <= 1,
                 CMD_ERR_ARG, "bad lattice description" );
    const int N = nx * ny * nz;
    *n = N;
    if ( !x )
        return CMD_OK;
    Mix rng{ seed };
    int k = 0;
    for ( int iz = 0; iz < nz; ++iz )
        for ( int iy = 0; iy < ny; ++iy )
            for ( int ix = 0; ix < nx; ++ix, ++k )
            {
                const int c[3] = { ix, iy, iz };
                for ( int d = 0; d < 3; ++d )
                    x[3 * k + d] =
                        ( c[d] + 0.5 + jitter * ( 2.0 * rng.uniform() - 1.0 ) ) * s;
                if ( id )
                    id[k] = k + 1;
            }
    if ( type )
    {
        const int n1 = (int)std::lround( N * frac1 );
        std::vector<int> t( N, 0 );
        for ( int i = 0; i < n1; ++i )
            t[i] = 1;
        for ( int i = N - 1; i > 0; --i )
        {
            const int j = (int)( rng.next() % (uint64_t)( i + 1 ) );
            std::swap( t[i], t[j] );
        }
        for ( int i = 0; i < N; ++i )
            type[i] = t[i];
    }
    return CMD_OK;
}

cmd_status cmd_force_nnp_model( const cmd_force *f, const cmd_nnp_model **m )
{
    CMD_REQUIRE( f && f->kind == 2 && m, CMD_ERR_ARG, "bad argument" );
    *m = &f->nnp->model;
    return CMD_OK;
}

// Work of the last compute under the flop model of roofline.h: out[0]
// neighbours inside rc, [1] triplets, [2] FP32 flops of the symmetry
// functions, [3] of the force chain, [4] of the network layers, [5] atoms.
cmd_status cmd_force_nnp_work( cmd_force *f, double out[6] )
{
    namespace F = cmd::nnp_flops;
    CMD_REQUIRE( f && f->kind == 2 && out, CMD_ERR_ARG, "bad argument" );
    cmd_nnp *M = f->nnp;
    double nb = 0.0, rad = 0.0, tri = 0.0, ang = 0.0, sf = 0.0, fc = 0.0;
    if ( M->n_last > 0 && M->terms.p )
    {
        std::vector<int4> t( M->n_last );
        CMD_HIP( hipMemcpy( t.data(), M->terms.p, sizeof( int4 ) * t.size(),
                            hipMemcpyDeviceToHost ) );
        // the kernels' path: the shorter one when every element's angular
        // rs are 0 (the launch in nnp_compute)
        bool rs0 = true;
        for ( int e = 0; e < M->plan.nelem; ++e )
            for ( int g = 0; g < M->plan.el[e].nag; ++g )
                rs0 = rs0 && M->plan.el[e].a_rs[g] == 0.f;
        for ( const int4 &a : t )
        {
            const int e = std::min( std::max( a.z, 0 ), M->plan.nelem - 1 );
            const cmd::NnpElemPlan &Q = M->plan.el[e];
            const double tr = a.y, angp = tr * Q.nap, angg = tr * Q.nag;
            nb += a.x;
            rad += (double)a.x * Q.nrg;
            tri += tr;
            ang += tr * Q.nag * Q.nap;
            sf += rs0 ? tr * F::SF_TRIP_RS0 + angp * F::SF_ANGP_RS0 + angg * F::SF_ANGG_RS0
                      : tr * F::SF_TRIP + angp * F::SF_ANGP + angg * F::SF_ANGG;
            fc += rs0 ? tr * F::F_TRIP_RS0 + angp * F::F_ANGP + angg * F::F_ANGG_RS0
                      : tr * F::F_TRIP + angp * F::F_ANGP + angg * F::F_ANGG;
        }
    }
    out[0] = nb;
    out[1] = tri;
    out[2] = nb * F::SF_NB + rad * F::SF_RAD + sf + ang * F::SF_ANG;
    out[3] = nb * F::F_NB + rad * F::F_RAD + fc + ang * F::F_ANG;
    // layers: the element with the largest network stands for all (the
    // models here give every element the same shape)
    double fwd = 0.0;
    for ( int e = 0; e < M->plan.nelem; ++e )
    {
        const cmd::NnpElemPlan &Q = M->plan.el[e];
        // nsf -> h1 -> h2 -> 1
        const double per = 2.0 * ( (double)Q.nsf * Q.h1 + (double)Q.h1 * Q.h2 + Q.h2 );
        fwd = std::max( fwd, per );
    }
    out[4] = F::MLP_PASSES * fwd * M->n_last;
    out[5] = M->n_last;
    return CMD_OK;
}

// Per-atom symmetry functions, dE/dG (both ld = stride floats per atom) and
// atomic energies of the last compute (any pointer may be NULL).
cmd_status cmd_force_nnp_atomic( cmd_force *f, cmd_system *s, int ld, float *G,
                                 float *dEdG, float *Ei )
{
    CMD_REQUIRE( f && f->kind == 2 && s && ld >= 1, CMD_ERR_ARG, "bad argument" );
    cmd_nnp *M = f->nnp;
    const int n = s->n_local;
    CMD_REQUIRE( M->cap >= n, CMD_ERR_STATE, "no NNP force computed yet" );
    std::vector<float> h( (size_t)std::max( n, 1 ) * NSF_LD );
    auto rows = [&]( const float *d, float *o ) -> cmd_status
    {
        CMD_HIP( hipMemcpy( h.data(), d, sizeof( float ) * (size_t)n * NSF_LD,
                            hipMemcpyDeviceToHost ) );
        for ( int i = 0; i < n; ++i )
            for ( int k = 0; k < std::min( ld, NSF_LD ); ++k )
                o[(size_t)i * ld + k] = h[(size_t)i * NSF_LD + k];
        return CMD_OK;
    };
    if ( G && n > 0 )
        CMD_TRY( rows( M->G.p, G ) );
    if ( dEdG && n > 0 )
        CMD_TRY( rows( M->dEdG.p, dEdG ) );
    if ( Ei && n > 0 )
        CMD_HIP( hipMemcpy( Ei, M->Ei.p, sizeof( float ) * n, hipMemcpyDeviceToHost ) );
    return CMD_OK;
}

} // extern "C"

namespace cmd
{
cmd_status nnp_model_copy( const cmd_nnp_model *m, cmd_nnp_model **out )
{
    *out = new cmd_nnp_model( *m );
    return CMD_OK;
}

// Atomic numbers of the element symbols n2p2 files name (weights.%03d.data).
static int atomic_number( const std::string &sym )
{
    static const char *const table[] = {
        "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
        "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr",
        "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
        "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
        "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd",
        "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
        "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po",
        "At", "Rn" };
    for ( int z = 0; z < (int)( sizeof( table ) / sizeof( table[0] ) ); ++z )
        if ( sym == table[z] )
            return z + 1;
    return 0;
}

cmd_status nnp_load_deck_model( const char *dir, cmd_nnp_model **out )
{
    CMD_REQUIRE( dir && dir[0], CMD_ERR_ARG,
                 "pair_style nnp needs a model directory (dir <path>)" );
    std::ifstream in( std::string( dir ) + "/input.nn" );
    CMD_REQUIRE( in.good(), CMD_ERR_IO, "cannot read %s/input.nn", dir );
    std::vector<std::string> names;
    std::string line;
    while ( std::getline( in, line ) )
    {
        const auto w = words_of( line );
        if ( w.size() >= 2 && w[0] == "elements" )
            names.assign( w.begin() + 1, w.end() );
    }
    CMD_REQUIRE( !names.empty() && (int)names.size() <= NNP_MAXEL, CMD_ERR_IO,
                 "input.nn must list 1..%d elements", NNP_MAXEL );
    std::vector<const char *> np;
    std::vector<int> z;
    for ( const auto &nm : names )
    {
        np.push_back( nm.c_str() );
        z.push_back( atomic_number( nm ) );
        CMD_REQUIRE( z.back() > 0, CMD_ERR_IO, "unknown element symbol %s",
                     nm.c_str() );
    }
    return cmd_nnp_model_load_n2p2( dir, (int)names.size(), np.data(), z.data(),
                                    out );
}

cmd_status nnp_destroy( cmd_nnp *M )
{
    delete M;
    return CMD_OK;
}
cmd_status nnp_set_precision( cmd_nnp *M, int prec )
{
    M->prec = prec;
    return CMD_OK;
}
// ForceNNP::compute_energy (impl.h:101-130): sum of atomic energies plus
// N_local times element 0's offset, over convEnergy (1: no normalisation),
// plus N times meanEnergy (0), times the hard-coded 27.211384021355236.
__global__ void k_nnp_ei_f64( int n, const float *__restrict__ Ei, double *__restrict__ out )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n )
        out[i] = (double)Ei[i];
}

cmd_status nnp_atom_energy( cmd_nnp *M, int n, double *out, hipStream_t st )
{
    if ( n <= 0 )
        return CMD_OK;
    CMD_REQUIRE( M->cap >= n, CMD_ERR_STATE, "no atomic energies for %d atoms", n );
    k_nnp_ei_f64<<<( n + 255 ) / 256, 256, 0, st>>>( n, M->Ei.p, out );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

double nnp_energy_offset( const cmd_nnp *M ) { return M->model.el[0].atom_energy; }

cmd_status nnp_reference_energy( cmd_nnp *M, cmd_system *S, double *pe )
{
    double sum = 0.0;
    CMD_TRY( nnp_energy_sum( M, S, &sum ) );
    *pe = ( sum + S->n_local * M->model.el[0].atom_energy ) * 27.211384021355236;
    return CMD_OK;
}
} // namespace cmd
