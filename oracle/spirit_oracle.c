/*
 * spirit_oracle.c -- CPU restatement of the reference's FP64 effective-field
 * and LLG step path.  TEST INFRASTRUCTURE ONLY (see spirit_oracle.h).
 *
 * Reference files followed (paths relative to the reference checkout):
 *   core/src/engine/Neighbours.cpp, core/src/data/Geometry.cpp,
 *   core/src/engine/Hamiltonian_Heisenberg.cpp, core/include/engine/Vectormath.hpp,
 *   core/src/engine/Vectormath.cpp, core/src/engine/Method_LLG.cpp,
 *   core/include/engine/Solver_{Depondt,Heun,SIB,RK4}.hpp, core/src/engine/Solver_Kernels.cpp
 *
 * Arithmetic follows the reference's expression order (Eigen evaluates small
 * fixed-size expressions component by component, left to right), compiled with
 * -ffp-contract=off so that no FMA is introduced.
 */
#include "spirit_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* small vector helpers (Eigen semantics)                              */
/* ------------------------------------------------------------------ */
static inline double v3_dot( const double * a, const double * b )
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double v3_sqnorm( const double * a )
{
    return a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
}
static inline void v3_cross( const double * a, const double * b, double * o )
{
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0]     = x;
    o[1]     = y;
    o[2]     = z;
}
/* Eigen normalize(): divide by sqrt(squaredNorm) if > 0, else unchanged (Dot.h:145-151) */
static inline void v3_normalize( double * a )
{
    double z = v3_sqnorm( a );
    if( z > 0 )
    {
        double n = sqrt( z );
        a[0] /= n;
        a[1] /= n;
        a[2] /= n;
    }
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */
struct orc_sys
{
    orc_params p;
    int nos;
    int n_cells_total;
    double * positions; /* AoS nos*3 */
    double * mu_s;      /* nos */
    double * site_K;    /* nos or NULL */
    /* DDI (FFT) */
    int padded[3];
    int M;
    int n_inter;
    int inter_lookup[ORC_MAX_BASIS * ORC_MAX_BASIS];
    double * d_hat;   /* complex [n_inter][6][M] interleaved */
    double * s_hat;   /* complex [n_atoms][3][M] */
    double * res_hat; /* complex [n_atoms][3][M] */
    /* DDI (cutoff): pairs within ddi_radius with their distances and unit vectors */
    int n_ddi_pairs;
    int * ddi_pairs;    /* [n][5]: i, j, da, db, dc */
    double * ddi_mag;   /* [n] */
    double * ddi_n;     /* [n][3] */
    /* scratch for LLG */
    double *f1, *fv1, *f2, *fv2, *sp, *tmp, *xi, *k1, *k2, *k3, *k4, *g;
};

/* positions, core/src/data/Geometry.cpp:142-161 */
static void compute_positions( const orc_params * p, double * pos )
{
    int Na = p->n_cells[0], Nb = p->n_cells[1], Nc = p->n_cells[2], N = p->n_cell_atoms;
    for( int dc = 0; dc < Nc; ++dc )
        for( int db = 0; db < Nb; ++db )
            for( int da = 0; da < Na; ++da )
                for( int iatom = 0; iatom < N; ++iatom )
                {
                    long ispin = iatom + (long)dc * N * Nb * Na + (long)db * N * Na + (long)da * N;
                    for( int c = 0; c < 3; ++c )
                    {
                        double v = ( (double)da + p->cell_atoms[iatom][0] ) * p->bravais[0][c]
                                   + ( (double)db + p->cell_atoms[iatom][1] ) * p->bravais[1][c]
                                   + ( (double)dc + p->cell_atoms[iatom][2] ) * p->bravais[2][c];
                        pos[3 * ispin + c] = p->lattice_constant * v;
                    }
                }
}

/* cell-0 position of basis atom i (positions[i] for i < n_cell_atoms) */
static void basis_position( const orc_params * p, int iatom, double * out )
{
    for( int c = 0; c < 3; ++c )
    {
        double v = ( 0.0 + p->cell_atoms[iatom][0] ) * p->bravais[0][c]
                   + ( 0.0 + p->cell_atoms[iatom][1] ) * p->bravais[1][c]
                   + ( 0.0 + p->cell_atoms[iatom][2] ) * p->bravais[2][c];
        out[c] = p->lattice_constant * v;
    }
}

static inline int imin( int a, int b )
{
    return a < b ? a : b;
}

/* core/src/engine/Neighbours.cpp:15-81 */
void orc_shell_radii( const orc_params * p, int n_shells, double * radii )
{
    const double min_shell_width = 1e-3;
    double ta[3], tb[3], tc[3];
    for( int c = 0; c < 3; ++c )
    {
        ta[c] = p->lattice_constant * p->bravais[0][c];
        tb[c] = p->lattice_constant * p->bravais[1][c];
        tc[c] = p->lattice_constant * p->bravais[2][c];
    }
    int max_n = n_shells + 2;
    int i_max = imin( max_n, p->n_cells[0] - 1 );
    int j_max = imin( max_n, p->n_cells[1] - 1 );
    int k_max = imin( max_n, p->n_cells[2] - 1 );
    if( sqrt( v3_sqnorm( ta ) ) == 0.0 )
        i_max = 0;
    if( sqrt( v3_sqnorm( tb ) ) == 0.0 )
        j_max = 0;
    if( sqrt( v3_sqnorm( tc ) ) == 0.0 )
        k_max = 0;

    double outermost = 0, previous = 0;
    for( int s = 0; s < n_shells; ++s )
    {
        radii[s]  = 0;
        previous  = outermost;
        outermost = DBL_MAX;
        for( int a1 = 0; a1 < p->n_cell_atoms; ++a1 )
        {
            const double * pos_one = p->cell_atoms[a1];
            for( int i = i_max; i >= 0; --i )
                for( int j = j_max; j >= -j_max; --j )
                    for( int k = k_max; k >= -k_max; --k )
                        for( int a2 = 0; a2 < p->n_cell_atoms; ++a2 )
                        {
                            if( !( a1 == a2 && i == 0 && j == 0 && k == 0 ) )
                            {
                                double d[3];
                                for( int c = 0; c < 3; ++c )
                                {
                                    double pos_two = p->cell_atoms[a2][c] + i * ta[c] + j * tb[c] + k * tc[c];
                                    d[c]           = pos_one[c] - pos_two;
                                }
                                double delta = sqrt( v3_sqnorm( d ) );
                                if( delta - previous > min_shell_width && delta < outermost )
                                {
                                    outermost = delta;
                                    radii[s]  = delta;
                                }
                            }
                        }
        }
    }
}

/* core/src/engine/Neighbours.cpp:83-156 */
int orc_neighbours_in_shells( const orc_params * p, int n_shells, int redundant, int * out, int max_pairs )
{
    const double min_shell_width = 1e-3;
    double radii[64];
    if( n_shells > 64 )
        return 0;
    orc_shell_radii( p, n_shells, radii );
    double ta[3], tb[3], tc[3];
    for( int c = 0; c < 3; ++c )
    {
        ta[c] = p->lattice_constant * p->bravais[0][c];
        tb[c] = p->lattice_constant * p->bravais[1][c];
        tc[c] = p->lattice_constant * p->bravais[2][c];
    }
    int max_n = n_shells + 2;
    int i_max = imin( max_n, p->n_cells[0] - 1 );
    int j_max = imin( max_n, p->n_cells[1] - 1 );
    int k_max = imin( max_n, p->n_cells[2] - 1 );
    int i_min = -i_max, j_min = -j_max, k_min = -k_max;
    if( sqrt( v3_sqnorm( ta ) ) == 0.0 )
        i_max = 0;
    if( sqrt( v3_sqnorm( tb ) ) == 0.0 )
        j_max = 0;
    if( sqrt( v3_sqnorm( tc ) ) == 0.0 )
        k_max = 0;

    int n = 0;
    for( int a1 = 0; a1 < p->n_cell_atoms; ++a1 )
    {
        int second_atom_min = redundant ? 0 : a1;
        const double * pos_one = p->cell_atoms[a1];
        for( int s = 0; s < n_shells; ++s )
        {
            double radius = radii[s];
            for( int i = i_max; i >= i_min; --i )
                for( int j = j_max; j >= j_min; --j )
                    for( int k = k_max; k >= k_min; --k )
                        for( int a2 = second_atom_min; a2 < p->n_cell_atoms; ++a2 )
                        {
                            if( ( a2 > a1 ) || ( i > 0 || ( i == 0 && j > 0 ) || ( i == 0 && j == 0 && k > 0 ) )
                                || redundant )
                            {
                                double d[3];
                                for( int c = 0; c < 3; ++c )
                                {
                                    double pos_two = p->cell_atoms[a2][c] + i * ta[c] + j * tb[c] + k * tc[c];
                                    d[c]           = pos_one[c] - pos_two;
                                }
                                double delta = sqrt( v3_sqnorm( d ) );
                                if( fabs( delta - radius ) < min_shell_width )
                                {
                                    if( n < max_pairs )
                                    {
                                        int * o = out + 6 * n;
                                        o[0]    = a1;
                                        o[1]    = a2;
                                        o[2]    = i;
                                        o[3]    = j;
                                        o[4]    = k;
                                        o[5]    = s;
                                    }
                                    ++n;
                                }
                            }
                        }
        }
    }
    return n <= max_pairs ? n : -n;
}

/* core/src/engine/Neighbours.cpp:249-286 */
void orc_dmi_normal( const orc_params * p, const int pair[5], int chirality, double out[3] )
{
    double ipos[3], jpos[3], d[3];
    basis_position( p, pair[0], ipos );
    basis_position( p, pair[1], jpos );
    for( int c = 0; c < 3; ++c )
    {
        double ta = p->lattice_constant * p->bravais[0][c];
        double tb = p->lattice_constant * p->bravais[1][c];
        double tc = p->lattice_constant * p->bravais[2][c];
        jpos[c]   = jpos[c] + pair[2] * ta + pair[3] * tb + pair[4] * tc;
    }
    const double ez[3] = { 0, 0, 1 };
    if( chirality == 1 || chirality == 2 || chirality == -2 )
    {
        for( int c = 0; c < 3; ++c )
            d[c] = jpos[c] - ipos[c];
        v3_normalize( d );
        if( chirality == 1 )
        {
            out[0] = d[0], out[1] = d[1], out[2] = d[2];
        }
        else if( chirality == 2 )
            v3_cross( d, ez, out );
        else
            v3_cross( ez, d, out );
    }
    else if( chirality == -1 )
    {
        for( int c = 0; c < 3; ++c )
            d[c] = ipos[c] - jpos[c];
        v3_normalize( d );
        out[0] = d[0], out[1] = d[1], out[2] = d[2];
    }
    else
    {
        out[0] = out[1] = out[2] = 0;
    }
}

/* Hamiltonian_Heisenberg.cpp:101-160 (shell branch, redundant neighbours) */
int orc_setup_shells( orc_params * p, int n_shells_exch, const double * jij, int n_shells_dmi, const double * dij,
                      int chirality )
{
    int buf[6 * ORC_MAX_PAIRS];
    p->n_exch = 0;
    p->n_dmi  = 0;
    if( n_shells_exch > 0 )
    {
        int n = orc_neighbours_in_shells( p, n_shells_exch, 1, buf, ORC_MAX_PAIRS );
        if( n < 0 )
            return -1;
        for( int k = 0; k < n; ++k )
        {
            for( int q = 0; q < 5; ++q )
                p->exch[k][q] = buf[6 * k + q];
            p->exch_J[k] = jij[buf[6 * k + 5]];
        }
        p->n_exch = n;
    }
    if( n_shells_dmi > 0 )
    {
        int n = orc_neighbours_in_shells( p, n_shells_dmi, 1, buf, ORC_MAX_PAIRS );
        if( n < 0 )
            return -1;
        for( int k = 0; k < n; ++k )
        {
            for( int q = 0; q < 5; ++q )
                p->dmi[k][q] = buf[6 * k + q];
            orc_dmi_normal( p, p->dmi[k], chirality, p->dmi_normal[k] );
            p->dmi_D[k] = dij[buf[6 * k + 5]];
        }
        p->n_dmi = n;
    }
    return 0;
}

/* Hamiltonian_Heisenberg.cpp:121-133 and 152-163 */
void orc_make_pairs_redundant( orc_params * p )
{
    int ne = p->n_exch;
    for( int k = 0; k < ne && p->n_exch < ORC_MAX_PAIRS; ++k )
    {
        int * q = p->exch[p->n_exch];
        q[0]    = p->exch[k][1];
        q[1]    = p->exch[k][0];
        q[2]    = -p->exch[k][2];
        q[3]    = -p->exch[k][3];
        q[4]    = -p->exch[k][4];
        p->exch_J[p->n_exch] = p->exch_J[k];
        p->n_exch++;
    }
    int nd = p->n_dmi;
    for( int k = 0; k < nd && p->n_dmi < ORC_MAX_PAIRS; ++k )
    {
        int * q = p->dmi[p->n_dmi];
        q[0]    = p->dmi[k][1];
        q[1]    = p->dmi[k][0];
        q[2]    = -p->dmi[k][2];
        q[3]    = -p->dmi[k][3];
        q[4]    = -p->dmi[k][4];
        p->dmi_D[p->n_dmi] = p->dmi_D[k];
        for( int c = 0; c < 3; ++c )
            p->dmi_normal[p->n_dmi][c] = -p->dmi_normal[k][c];
        p->n_dmi++;
    }
}

/* core/include/engine/Vectormath.hpp:437-528 */
int orc_idx_from_pair( int ispin, const int bc[3], const int n_cells[3], int N, const int pair[5] )
{
    if( pair[0] != ispin % N )
        return -1;
    int Na = n_cells[0], Nb = n_cells[1], Nc = n_cells[2];
    if( abs( pair[2] ) > Na || abs( pair[3] ) > Nb || abs( pair[4] ) > Nc )
        return -1;
    int nic = ispin / ( N * Na * Nb );
    int nib = ( ispin - nic * N * Na * Nb ) / ( N * Na );
    int nia = ( ispin - nic * N * Na * Nb - nib * N * Na ) / N;
    int nja = nia + pair[2], njb = nib + pair[3], njc = nic + pair[4];
    if( bc[0] || ( 0 <= nja && nja < Na ) )
    {
        if( nja < 0 )
            nja += Na;
        if( nja >= Na )
            nja -= Na;
    }
    else
        return -1;
    if( bc[1] || ( 0 <= njb && njb < Nb ) )
    {
        if( njb < 0 )
            njb += Nb;
        if( njb >= Nb )
            njb -= Nb;
    }
    else
        return -1;
    if( bc[2] || ( 0 <= njc && njc < Nc ) )
    {
        if( njc < 0 )
            njc += Nc;
        if( njc >= Nc )
            njc -= Nc;
    }
    else
        return -1;
    return pair[1] + nja * N + njb * N * Na + njc * N * Na * Nb;
}

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed-radix DFT (decimation in time), any length      */
/* ------------------------------------------------------------------ */
typedef struct
{
    double re, im;
} cpx;

static int smallest_factor( int n )
{
    if( n % 4 == 0 )
        return 4;
    if( n % 2 == 0 )
        return 2;
    for( int f = 3; f * f <= n; f += 2 )
        if( n % f == 0 )
            return f;
    return n;
}

/* W: table of the N-th roots exp(sign*2*pi*i*e/N), e in [0,N).  wstep = N/n */
static void fft_rec( const cpx * in, long istride, cpx * out, int n, const cpx * W, int N, int wstep, cpx * scratch )
{
    if( n == 1 )
    {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor( n );
    int m = n / p;
    for( int r = 0; r < p; ++r )
        fft_rec( in + r * istride, istride * p, out + r * m, m, W, N, wstep * p, scratch );
    for( int k = 0; k < m; ++k )
    {
        for( int q = 0; q < p; ++q )
        {
            double sr = 0, si = 0;
            for( int r = 0; r < p; ++r )
            {
                long e      = ( (long)r * ( k + (long)q * m ) ) % n;
                const cpx w = W[e * wstep];
                const cpx y = out[r * m + k];
                sr += y.re * w.re - y.im * w.im;
                si += y.re * w.im + y.im * w.re;
            }
            scratch[q].re = sr;
            scratch[q].im = si;
        }
        for( int q = 0; q < p; ++q )
            out[k + q * m] = scratch[q];
    }
}

static cpx * make_roots( int N, int sign )
{
    cpx * W = (cpx *)malloc( sizeof( cpx ) * (size_t)N );
    for( int e = 0; e < N; ++e )
    {
        long double a = (long double)sign * 2.0L * 3.14159265358979323846264338327950288L * (long double)e
                        / (long double)N;
        W[e].re = (double)cosl( a );
        W[e].im = (double)sinl( a );
    }
    return W;
}

void orc_fft( double * data, int n, int sign )
{
    cpx * W   = make_roots( n, sign );
    cpx * in  = (cpx *)malloc( sizeof( cpx ) * (size_t)n );
    cpx * scr = (cpx *)malloc( sizeof( cpx ) * (size_t)n );
    memcpy( in, data, sizeof( cpx ) * (size_t)n );
    fft_rec( in, 1, (cpx *)data, n, W, n, 1, scr );
    free( W );
    free( in );
    free( scr );
}

/* Whether the lines of a 3-D transform are spread over the OpenMP threads (1, the default)
   or run on one thread, as the reference's kissFFT build transforms (kiss_fftndr runs each
   transform serially; the reference's batch_Four_3D loops over the transforms in turn). */
static int g_fft_parallel = 1;

void orc_set_fft_parallel( int on )
{
    g_fft_parallel = on ? 1 : 0;
}

/* 3-D complex FFT over grid [c][b][a] (a fastest), `count` consecutive arrays of M points. */
static void fft3d( cpx * data, int count, const int dims[3], int sign )
{
    int Pa = dims[0], Pb = dims[1], Pc = dims[2];
    long M = (long)Pa * Pb * Pc;
    for( int axis = 0; axis < 3; ++axis )
    {
        int n = dims[axis];
        if( n == 1 )
            continue;
        cpx * W      = make_roots( n, sign );
        long stride  = axis == 0 ? 1 : ( axis == 1 ? Pa : (long)Pa * Pb );
        long n_lines = M / n;
#pragma omp parallel if( g_fft_parallel )
        {
            cpx * buf = (cpx *)malloc( sizeof( cpx ) * (size_t)n * 3 );
            cpx * out = buf + n;
            cpx * scr = buf + 2 * n;
#pragma omp for schedule( static )
            for( long t = 0; t < (long)count * n_lines; ++t )
            {
                long arr  = t / n_lines;
                long line = t % n_lines;
                long base;
                if( axis == 0 )
                    base = line * Pa;
                else if( axis == 1 )
                    base = ( line / Pa ) * (long)Pa * Pb + ( line % Pa );
                else
                    base = line;
                cpx * d = data + arr * M + base;
                for( int i = 0; i < n; ++i )
                    buf[i] = d[i * stride];
                fft_rec( buf, 1, out, n, W, n, 1, scr );
                for( int i = 0; i < n; ++i )
                    d[i * stride] = out[i];
            }
            free( buf );
        }
        free( W );
    }
}

/* ------------------------------------------------------------------ */
/* DDI setup: Hamiltonian_Heisenberg.cpp:1406-1595                      */
/* ------------------------------------------------------------------ */
static double ddi_mult( void )
{
    return ORC_MU_0 * ORC_MU_B * ORC_MU_B / ( 4 * ORC_PI * 1e-30 );
}

void orc_ddi_padded( const orc_sys * s, int out[3] )
{
    for( int i = 0; i < 3; ++i )
    {
        out[i]   = s->p.n_cells[i];
        int pad  = s->p.n_cells[i] > 1 && ( s->p.bc[i] == 0 || s->p.ddi_pb_zero_padding );
        if( pad )
            out[i] *= 2;
    }
}

static void ddi_inter_lookup( const orc_params * p, int * lookup, int * n_inter )
{
    int b_inter = -1;
    int N       = p->n_cell_atoms;
    for( int b1 = 0; b1 < N; ++b1 )
        for( int b2 = 0; b2 < N; ++b2 )
        {
            if( b1 == b2 && b1 != 0 )
            {
                lookup[b1 + b2 * N] = 0;
                continue;
            }
            b_inter++;
            lookup[b1 + b2 * N] = b_inter;
        }
    *n_inter = b_inter + 1;
}

/* Real-space dipole matrices on the padded grid, Hamiltonian_Heisenberg.cpp:1406-1499.
   out: [b_inter][6][M], grid index a + Pa*(b + Pb*c). */
void orc_ddi_dipole_matrices( const orc_sys * s, double * out )
{
    const orc_params * p = &s->p;
    const double mult    = ddi_mult();
    int Na = p->n_cells[0], Nb = p->n_cells[1], Nc = p->n_cells[2];
    int Pa = s->padded[0], Pb = s->padded[1], Pc = s->padded[2];
    long M    = (long)Pa * Pb * Pc;
    int img_a = p->bc[0] == 0 ? 0 : p->ddi_n_periodic_images[0];
    int img_b = p->bc[1] == 0 ? 0 : p->ddi_n_periodic_images[1];
    int img_c = p->bc[2] == 0 ? 0 : p->ddi_n_periodic_images[2];
    int N     = p->n_cell_atoms;
    int b_inter = -1;
    for( int b1 = 0; b1 < N; ++b1 )
        for( int b2 = 0; b2 < N; ++b2 )
        {
            if( b1 == b2 && b1 != 0 )
                continue;
            b_inter++;
            double * o = out + (long)b_inter * 6 * M;
#pragma omp parallel for collapse( 2 )
            for( int c = 0; c < Pc; ++c )
                for( int b = 0; b < Pb; ++b )
                    for( int a = 0; a < Pa; ++a )
                    {
                        int a_idx = a < Na ? a : a - Pa;
                        int b_idx = b < Nb ? b : b - Pb;
                        int c_idx = c < Nc ? c : c - Pc;
                        double Dxx = 0, Dxy = 0, Dxz = 0, Dyy = 0, Dyz = 0, Dzz = 0;
                        for( int a_pb = -img_a; a_pb <= img_a; a_pb++ )
                            for( int b_pb = -img_b; b_pb <= img_b; b_pb++ )
                                for( int c_pb = -img_c; c_pb <= img_c; c_pb++ )
                                {
                                    double fa = (double)( a_idx + a_pb * Na ) + p->cell_atoms[b1][0]
                                                - p->cell_atoms[b2][0];
                                    double fb = (double)( b_idx + b_pb * Nb ) + p->cell_atoms[b1][1]
                                                - p->cell_atoms[b2][1];
                                    double fc = (double)( c_idx + c_pb * Nc ) + p->cell_atoms[b1][2]
                                                - p->cell_atoms[b2][2];
                                    double diff[3];
                                    for( int q = 0; q < 3; ++q )
                                        diff[q] = p->lattice_constant
                                                  * ( fa * p->bravais[0][q] + fb * p->bravais[1][q]
                                                      + fc * p->bravais[2][q] );
                                    double d = sqrt( v3_sqnorm( diff ) );
                                    if( d > 1e-10 )
                                    {
                                        double d3 = d * d * d;
                                        double d5 = d * d * d * d * d;
                                        Dxx += mult * ( 3 * diff[0] * diff[0] / d5 - 1 / d3 );
                                        Dxy += mult * 3 * diff[0] * diff[1] / d5;
                                        Dxz += mult * 3 * diff[0] * diff[2] / d5;
                                        Dyy += mult * ( 3 * diff[1] * diff[1] / d5 - 1 / d3 );
                                        Dyz += mult * 3 * diff[1] * diff[2] / d5;
                                        Dzz += mult * ( 3 * diff[2] * diff[2] / d5 - 1 / d3 );
                                    }
                                }
                        long idx       = a + (long)Pa * ( b + (long)Pb * c );
                        o[0 * M + idx] = Dxx;
                        o[1 * M + idx] = Dxy;
                        o[2 * M + idx] = Dxz;
                        o[3 * M + idx] = Dyy;
                        o[4 * M + idx] = Dyz;
                        o[5 * M + idx] = Dzz;
                    }
        }
}

static void ddi_prepare( orc_sys * s )
{
    orc_ddi_padded( s, s->padded );
    s->M = s->padded[0] * s->padded[1] * s->padded[2];
    ddi_inter_lookup( &s->p, s->inter_lookup, &s->n_inter );
    long M     = s->M;
    double * D = (double *)malloc( sizeof( double ) * 6 * (size_t)s->n_inter * (size_t)M );
    orc_ddi_dipole_matrices( s, D );
    s->d_hat = (double *)calloc( 2 * 6 * (size_t)s->n_inter * (size_t)M, sizeof( double ) );
    for( long t = 0; t < 6L * s->n_inter * M; ++t )
        s->d_hat[2 * t] = D[t];
    free( D );
    fft3d( (cpx *)s->d_hat, 6 * s->n_inter, s->padded, -1 );
    s->s_hat   = (double *)malloc( sizeof( double ) * 2 * 3 * (size_t)s->p.n_cell_atoms * (size_t)M );
    s->res_hat = (double *)malloc( sizeof( double ) * 2 * 3 * (size_t)s->p.n_cell_atoms * (size_t)M );
}

/* Gradient_DDI_FFT + FFT_Spins, Hamiltonian_Heisenberg.cpp:920-1014, 1373-1404 */
static void gradient_ddi_fft( orc_sys * s, const double * spins, double * grad )
{
    const orc_params * p = &s->p;
    int Na = p->n_cells[0], Nb = p->n_cells[1], Nc = p->n_cells[2], N = p->n_cell_atoms;
    int Pa = s->padded[0], Pb = s->padded[1];
    long M = s->M;
    memset( s->s_hat, 0, sizeof( double ) * 2 * 3 * (size_t)N * (size_t)M );
#pragma omp parallel for collapse( 2 )
    for( int c = 0; c < Nc; ++c )
        for( int b = 0; b < Nb; ++b )
            for( int a = 0; a < Na; ++a )
                for( int bi = 0; bi < N; ++bi )
                {
                    long io  = bi + (long)N * ( a + (long)Na * ( b + (long)Nb * c ) );
                    long idx = a + (long)Pa * ( b + (long)Pb * c );
                    for( int q = 0; q < 3; ++q )
                        s->s_hat[2 * ( ( (long)bi * 3 + q ) * M + idx )] = spins[3 * io + q] * s->mu_s[io];
                }
    fft3d( (cpx *)s->s_hat, 3 * N, s->padded, -1 );

    const cpx * Sh = (const cpx *)s->s_hat;
    const cpx * Dh = (const cpx *)s->d_hat;
    cpx * R        = (cpx *)s->res_hat;
#pragma omp parallel for
    for( long k = 0; k < M; ++k )
    {
        for( int b1 = 0; b1 < N; ++b1 )
        {
            cpx acc[3] = { { 0, 0 }, { 0, 0 }, { 0, 0 } };
            for( int b2 = 0; b2 < N; ++b2 )
            {
                int bi         = s->inter_lookup[b1 + b2 * N];
                const cpx * d  = Dh + (long)bi * 6 * M + k;
                const cpx sx = Sh[( (long)b2 * 3 + 0 ) * M + k];
                const cpx sy = Sh[( (long)b2 * 3 + 1 ) * M + k];
                const cpx sz = Sh[( (long)b2 * 3 + 2 ) * M + k];
                /* rows (xx,xy,xz),(xy,yy,yz),(xz,yz,zz), FFT.hpp:89-96 */
                const int rows[3][3] = { { 0, 1, 2 }, { 1, 3, 4 }, { 2, 4, 5 } };
                for( int q = 0; q < 3; ++q )
                {
                    cpx d1 = d[rows[q][0] * M], d2 = d[rows[q][1] * M], d3 = d[rows[q][2] * M];
                    double re = d1.re * sx.re - d1.im * sx.im + d2.re * sy.re - d2.im * sy.im + d3.re * sz.re
                                - d3.im * sz.im;
                    double im = d1.re * sx.im + d1.im * sx.re + d2.re * sy.im + d2.im * sy.re + d3.re * sz.im
                                + d3.im * sz.re;
                    if( b2 == 0 )
                    {
                        acc[q].re = re;
                        acc[q].im = im;
                    }
                    else
                    {
                        acc[q].re += re;
                        acc[q].im += im;
                    }
                }
            }
            for( int q = 0; q < 3; ++q )
                R[( (long)b1 * 3 + q ) * M + k] = acc[q];
        }
    }
    fft3d( R, 3 * N, s->padded, +1 );

#pragma omp parallel for collapse( 2 )
    for( int c = 0; c < Nc; ++c )
        for( int b = 0; b < Nb; ++b )
            for( int a = 0; a < Na; ++a )
                for( int b1 = 0; b1 < N; ++b1 )
                {
                    long io  = b1 + (long)N * ( a + (long)Na * ( b + (long)Nb * c ) );
                    long idx = a + (long)Pa * ( b + (long)Pb * c );
                    for( int q = 0; q < 3; ++q )
                        grad[3 * io + q] -= s->mu_s[io] * R[( (long)b1 * 3 + q ) * M + idx].re / (double)M;
                }
}

/* Gradient_DDI_Direct, Hamiltonian_Heisenberg.cpp:1016-1071 */
static void gradient_ddi_direct( orc_sys * s, const double * spins, double * grad )
{
    const orc_params * p = &s->p;
    const double mult    = ddi_mult();
    int img_a = p->bc[0] == 0 ? 0 : p->ddi_n_periodic_images[0];
    int img_b = p->bc[1] == 0 ? 0 : p->ddi_n_periodic_images[1];
    int img_c = p->bc[2] == 0 ? 0 : p->ddi_n_periodic_images[2];
    int nos   = s->nos;
#pragma omp parallel for
    for( int i1 = 0; i1 < nos; i1++ )
    {
        for( int i2 = 0; i2 < nos; i2++ )
        {
            const double * m2 = spins + 3 * i2;
            double diff[3];
            for( int q = 0; q < 3; ++q )
                diff[q] = s->positions[3 * i2 + q] - s->positions[3 * i1 + q];
            double Dxx = 0, Dxy = 0, Dxz = 0, Dyy = 0, Dyz = 0, Dzz = 0;
            for( int a_pb = -img_a; a_pb <= img_a; a_pb++ )
                for( int b_pb = -img_b; b_pb <= img_b; b_pb++ )
                    for( int c_pb = -img_c; c_pb <= img_c; c_pb++ )
                    {
                        double di[3];
                        for( int q = 0; q < 3; ++q )
                            di[q] = diff[q] + a_pb * p->n_cells[0] * p->bravais[0][q] * p->lattice_constant
                                    + b_pb * p->n_cells[1] * p->bravais[1][q] * p->lattice_constant
                                    + c_pb * p->n_cells[2] * p->bravais[2][q] * p->lattice_constant;
                        double d = sqrt( v3_sqnorm( di ) );
                        if( d > 1e-10 )
                        {
                            double d3 = d * d * d;
                            double d5 = d * d * d * d * d;
                            Dxx += mult * ( 3 * di[0] * di[0] / d5 - 1 / d3 );
                            Dxy += mult * 3 * di[0] * di[1] / d5;
                            Dxz += mult * 3 * di[0] * di[2] / d5;
                            Dyy += mult * ( 3 * di[1] * di[1] / d5 - 1 / d3 );
                            Dyz += mult * 3 * di[1] * di[2] / d5;
                            Dzz += mult * ( 3 * di[2] * di[2] / d5 - 1 / d3 );
                        }
                    }
            double mm = s->mu_s[i1] * s->mu_s[i2];
            grad[3 * i1 + 0] -= ( Dxx * m2[0] + Dxy * m2[1] + Dxz * m2[2] ) * mm;
            grad[3 * i1 + 1] -= ( Dxy * m2[0] + Dyy * m2[1] + Dyz * m2[2] ) * mm;
            grad[3 * i1 + 2] -= ( Dxz * m2[0] + Dyz * m2[1] + Dzz * m2[2] ) * mm;
        }
    }
}

/* Neighbours::Get_Pairs_in_Radius, Neighbours.cpp:158-246, and DDI_from_Pair, :288-307:
   every (iatom, jatom, translation) with 0 < |r_j + t - r_i| < radius, iatom outer, then the
   translations a, b, c ascending, jatom inner; the translation ranges from the lattice's
   bounding box (Geometry::calculateBounds, Geometry.cpp:729-744, which starts at the origin). */
static void ddi_cutoff_prepare( orc_sys * s )
{
    const orc_params * p = &s->p;
    const double radius  = p->ddi_radius;
    s->n_ddi_pairs       = 0;
    if( fabs( radius ) < 1e-6 )
        return;
    double bmin[3] = { 0, 0, 0 }, bmax[3] = { 0, 0, 0 };
    for( int i = 0; i < s->nos; ++i )
        for( int q = 0; q < 3; ++q )
        {
            double v = s->positions[3 * i + q];
            if( v < bmin[q] )
                bmin[q] = v;
            if( v > bmax[q] )
                bmax[q] = v;
        }
    int mx[3] = { 0, 0, 0 };
    double tv[3][3];
    for( int k = 0; k < 3; ++k )
    {
        for( int q = 0; q < 3; ++q )
            tv[k][q] = p->lattice_constant * p->bravais[k][q];
        double diff = bmax[k] - bmin[k];
        if( radius > 0 )
        {
            if( diff > 0 )
            {
                int m = (int)( 1.1 * radius * p->n_cells[k] / diff );
                mx[k] = p->n_cells[k] - 1 < m ? p->n_cells[k] - 1 : m;
            }
        }
        else
            mx[k] = p->n_cells[k] - 1;
        if( sqrt( v3_sqnorm( tv[k] ) ) == 0.0 )
            mx[k] = 0;
    }
    long cap = (long)p->n_cell_atoms * p->n_cell_atoms * ( 2 * mx[0] + 1 ) * ( 2 * mx[1] + 1 ) * ( 2 * mx[2] + 1 );
    s->ddi_pairs = (int *)malloc( sizeof( int ) * 5 * (size_t)cap );
    s->ddi_mag   = (double *)malloc( sizeof( double ) * (size_t)cap );
    s->ddi_n     = (double *)malloc( sizeof( double ) * 3 * (size_t)cap );
    int n        = 0;
    for( int ia = 0; ia < p->n_cell_atoms; ++ia )
    {
        double pi_[3];
        basis_position( p, ia, pi_ );
        for( int i = -mx[0]; i <= mx[0]; ++i )
            for( int j = -mx[1]; j <= mx[1]; ++j )
                for( int k = -mx[2]; k <= mx[2]; ++k )
                    for( int ja = 0; ja < p->n_cell_atoms; ++ja )
                    {
                        double pj[3], d[3];
                        basis_position( p, ja, pj );
                        for( int q = 0; q < 3; ++q )
                            pj[q] = ( ( pj[q] + i * tv[0][q] ) + j * tv[1][q] ) + k * tv[2][q];
                        for( int q = 0; q < 3; ++q )
                            d[q] = pi_[q] - pj[q];
                        double dist = sqrt( v3_sqnorm( d ) );
                        if( dist < radius && dist > 2.220446049250313e-16 )
                        {
                            int * e = s->ddi_pairs + 5 * n;
                            e[0] = ia, e[1] = ja, e[2] = i, e[3] = j, e[4] = k;
                            /* DDI_from_Pair: r_ij = r_j + t - r_i, its norm and direction */
                            double r[3];
                            for( int q = 0; q < 3; ++q )
                                r[q] = pj[q] - pi_[q];
                            double m      = sqrt( v3_sqnorm( r ) );
                            s->ddi_mag[n] = m;
                            for( int q = 0; q < 3; ++q )
                                s->ddi_n[3 * n + q] = m > 0 ? r[q] / m : 0.0;
                            ++n;
                        }
                    }
    }
    s->n_ddi_pairs = n;
}

/* Gradient_DDI_Cutoff, Hamiltonian_Heisenberg.cpp:880-918 (pairs in pair order, cells inner) */
static void gradient_ddi_cutoff( orc_sys * s, const double * spins, double * grad )
{
    const orc_params * p = &s->p;
    const double mult    = ddi_mult();
    for( int ip = 0; ip < s->n_ddi_pairs; ++ip )
    {
        if( !( s->ddi_mag[ip] > 0.0 ) )
            continue;
        const double c   = mult / pow( s->ddi_mag[ip], 3.0 );
        const int * pr   = s->ddi_pairs + 5 * ip;
        const double * n = s->ddi_n + 3 * ip;
#pragma omp parallel for
        for( int icell = 0; icell < s->n_cells_total; ++icell )
        {
            int i = pr[0] + icell * p->n_cell_atoms;
            int j = orc_idx_from_pair( i, p->bc, p->n_cells, p->n_cell_atoms, pr );
            if( j >= 0 )
            {
                const double * sj = spins + 3 * j;
                double sn         = v3_dot( sj, n );
                double f          = s->mu_s[j] * s->mu_s[i] * c;
                for( int q = 0; q < 3; ++q )
                    grad[3 * i + q] -= f * ( 3 * n[q] * sn - sj[q] );
            }
        }
    }
}

/* Hamiltonian_Heisenberg::Hessian, Hamiltonian_Heisenberg.cpp:1119-1245: the dense 3N x 3N
   embedding Hessian (row-major; index 3*ispin + component).  Anisotropy: every (alpha, beta)
   entry of -2 K n n^T lands on (3i+alpha, 3i+alpha) (:1133-1144), so the diagonal holds the
   row sums; exchange (:1150-1172) and DMI (:1174-1207) per redundant pair at (i, j); the
   dipolar block (:1210-1245) as the direct sum over the positions without periodic images,
   for any DDI method.  The per-site K_i extension is treated as a uniaxial term. */
void orc_hessian( orc_sys * s, const double * spins, double * H )
{
    (void)spins;
    const orc_params * p = &s->p;
    const long n3        = 3L * s->nos;
    const int N          = p->n_cell_atoms;
    memset( H, 0, sizeof( double ) * (size_t)n3 * (size_t)n3 );
    for( int icell = 0; icell < s->n_cells_total; ++icell )
    {
        for( int k = 0; k < p->n_aniso; ++k )
        {
            long ispin = (long)icell * N + p->aniso_idx[k];
            for( int a = 0; a < 3; ++a )
                for( int b = 0; b < 3; ++b )
                {
                    long r = 3 * ispin + a;
                    H[r * n3 + r] += -2.0 * p->aniso_K[k] * p->aniso_normal[k][a] * p->aniso_normal[k][b];
                }
        }
        if( s->site_K )
            for( int ib = 0; ib < N; ++ib )
            {
                long ispin = (long)icell * N + ib;
                for( int a = 0; a < 3; ++a )
                    for( int b = 0; b < 3; ++b )
                    {
                        long r = 3 * ispin + a;
                        H[r * n3 + r] += -2.0 * s->site_K[ispin] * p->site_aniso_normal[a] * p->site_aniso_normal[b];
                    }
            }
    }
    for( int icell = 0; icell < s->n_cells_total; ++icell )
        for( int k = 0; k < p->n_exch; ++k )
        {
            int ispin = p->exch[k][0] + icell * N;
            int jspin = orc_idx_from_pair( ispin, p->bc, p->n_cells, N, p->exch[k] );
            if( jspin >= 0 )
                for( int a = 0; a < 3; ++a )
                    H[( 3L * ispin + a ) * n3 + 3L * jspin + a] += -p->exch_J[k];
        }
    for( int icell = 0; icell < s->n_cells_total; ++icell )
        for( int k = 0; k < p->n_dmi; ++k )
        {
            int ispin = p->dmi[k][0] + icell * N;
            int jspin = orc_idx_from_pair( ispin, p->bc, p->n_cells, N, p->dmi[k] );
            if( jspin >= 0 )
            {
                long i = 3L * ispin, j = 3L * jspin;
                double D = p->dmi_D[k];
                const double * n = p->dmi_normal[k];
                H[( i + 2 ) * n3 + j + 1] += D * n[0];
                H[( i + 1 ) * n3 + j + 2] += -D * n[0];
                H[( i ) * n3 + j + 2] += D * n[1];
                H[( i + 2 ) * n3 + j] += -D * n[1];
                H[( i + 1 ) * n3 + j] += D * n[2];
                H[( i ) * n3 + j + 1] += -D * n[2];
            }
        }
    if( p->ddi_method != ORC_DDI_NONE )
    {
        const double mult = ddi_mult();
#pragma omp parallel for
        for( int i1 = 0; i1 < s->nos; ++i1 )
            for( int i2 = 0; i2 < s->nos; ++i2 )
            {
                double diff[3];
                for( int q = 0; q < 3; ++q )
                    diff[q] = s->positions[3 * i2 + q] - s->positions[3 * i1 + q];
                double d = sqrt( v3_sqnorm( diff ) );
                double Dxx = 0, Dxy = 0, Dxz = 0, Dyy = 0, Dyz = 0, Dzz = 0;
                if( d > 1e-10 )
                {
                    double d3 = d * d * d, d5 = d * d * d * d * d;
                    Dxx += mult * ( 3 * diff[0] * diff[0] / d5 - 1 / d3 );
                    Dxy += mult * 3 * diff[0] * diff[1] / d5;
                    Dxz += mult * 3 * diff[0] * diff[2] / d5;
                    Dyy += mult * ( 3 * diff[1] * diff[1] / d5 - 1 / d3 );
                    Dyz += mult * 3 * diff[1] * diff[2] / d5;
                    Dzz += mult * ( 3 * diff[2] * diff[2] / d5 - 1 / d3 );
                }
                const double m = -s->mu_s[i1] * s->mu_s[i2];
                const double T[3][3] = { { Dxx, Dxy, Dxz }, { Dxy, Dyy, Dyz }, { Dxz, Dyz, Dzz } };
                for( int a = 0; a < 3; ++a )
                    for( int b = 0; b < 3; ++b )
                        H[( 3L * i1 + a ) * n3 + 3L * i2 + b] += m * T[a][b];
            }
    }
}

int orc_ddi_cutoff_pairs( const orc_sys * s, int * pairs, double * magnitudes, double * normals, int max_pairs )
{
    for( int k = 0; k < s->n_ddi_pairs && k < max_pairs; ++k )
    {
        if( pairs )
            for( int q = 0; q < 5; ++q )
                pairs[5 * k + q] = s->ddi_pairs[5 * k + q];
        if( magnitudes )
            magnitudes[k] = s->ddi_mag[k];
        if( normals )
            for( int q = 0; q < 3; ++q )
                normals[3 * k + q] = s->ddi_n[3 * k + q];
    }
    return s->n_ddi_pairs;
}

/* ------------------------------------------------------------------ */
/* system                                                              */
/* ------------------------------------------------------------------ */
orc_sys * orc_create( const orc_params * p )
{
    orc_sys * s = (orc_sys *)calloc( 1, sizeof( orc_sys ) );
    s->p        = *p;
    s->nos      = p->n_cell_atoms * p->n_cells[0] * p->n_cells[1] * p->n_cells[2];
    s->n_cells_total = p->n_cells[0] * p->n_cells[1] * p->n_cells[2];
    s->positions     = (double *)malloc( sizeof( double ) * 3 * (size_t)s->nos );
    compute_positions( p, s->positions );
    s->mu_s = (double *)malloc( sizeof( double ) * (size_t)s->nos );
    for( int i = 0; i < s->nos; ++i )
        s->mu_s[i] = p->mu_s[i % p->n_cell_atoms];
    if( p->site_aniso_K )
    {
        s->site_K = (double *)malloc( sizeof( double ) * (size_t)s->nos );
        memcpy( s->site_K, p->site_aniso_K, sizeof( double ) * (size_t)s->nos );
    }
    s->p.site_aniso_K = NULL;
    if( p->ddi_method == ORC_DDI_FFT )
        ddi_prepare( s );
    else if( p->ddi_method == ORC_DDI_CUTOFF && p->ddi_radius >= 0 )
        ddi_cutoff_prepare( s );
    size_t n3 = sizeof( double ) * 3 * (size_t)s->nos;
    s->f1     = (double *)malloc( n3 );
    s->fv1    = (double *)malloc( n3 );
    s->f2     = (double *)malloc( n3 );
    s->fv2    = (double *)malloc( n3 );
    s->sp     = (double *)malloc( n3 );
    s->tmp    = (double *)malloc( n3 );
    s->xi     = (double *)calloc( 3 * (size_t)s->nos, sizeof( double ) );
    s->k1     = (double *)malloc( n3 );
    s->k2     = (double *)malloc( n3 );
    s->k3     = (double *)malloc( n3 );
    s->k4     = (double *)malloc( n3 );
    s->g      = (double *)malloc( n3 );
    return s;
}

void orc_destroy( orc_sys * s )
{
    if( !s )
        return;
    free( s->positions );
    free( s->mu_s );
    free( s->site_K );
    free( s->d_hat );
    free( s->s_hat );
    free( s->res_hat );
    free( s->ddi_pairs );
    free( s->ddi_mag );
    free( s->ddi_n );
    free( s->f1 );
    free( s->fv1 );
    free( s->f2 );
    free( s->fv2 );
    free( s->sp );
    free( s->tmp );
    free( s->xi );
    free( s->k1 );
    free( s->k2 );
    free( s->k3 );
    free( s->k4 );
    free( s->g );
    free( s );
}

int orc_nos( const orc_sys * s )
{
    return s->nos;
}

void orc_positions( const orc_sys * s, double * out )
{
    memcpy( out, s->positions, sizeof( double ) * 3 * (size_t)s->nos );
}

/* Gradient_Zeeman, Hamiltonian_Heisenberg.cpp:768-783 */
static void grad_zeeman( orc_sys * s, double * g )
{
    const double B = s->p.field_tesla * ORC_MU_B;
#pragma omp parallel for
    for( int i = 0; i < s->nos; ++i )
        for( int q = 0; q < 3; ++q )
            g[3 * i + q] -= s->mu_s[i] * B * s->p.field_normal[q];
}

/* Gradient_Anisotropy, :785-800 (+ per-site extension) */
static void grad_aniso( orc_sys * s, const double * sp, double * g )
{
    const orc_params * p = &s->p;
    int N                = p->n_cell_atoms;
#pragma omp parallel for
    for( int icell = 0; icell < s->n_cells_total; ++icell )
    {
        for( int ia = 0; ia < p->n_aniso; ++ia )
        {
            int i          = icell * N + p->aniso_idx[ia];
            const double * n = p->aniso_normal[ia];
            double d       = v3_dot( n, sp + 3 * i );
            for( int q = 0; q < 3; ++q )
                g[3 * i + q] -= 2.0 * p->aniso_K[ia] * n[q] * d;
        }
        if( s->site_K )
        {
            for( int b = 0; b < N; ++b )
            {
                int i          = icell * N + b;
                const double * n = p->site_aniso_normal;
                double d       = v3_dot( n, sp + 3 * i );
                for( int q = 0; q < 3; ++q )
                    g[3 * i + q] -= 2.0 * s->site_K[i] * n[q] * d;
            }
        }
    }
}

/* Gradient_Cubic_Anisotropy, :802-820 */
static void grad_cubic( orc_sys * s, const double * sp, double * g )
{
    const orc_params * p = &s->p;
    int N                = p->n_cell_atoms;
#pragma omp parallel for
    for( int icell = 0; icell < s->n_cells_total; ++icell )
        for( int ia = 0; ia < p->n_cubic; ++ia )
        {
            int i = icell * N + p->cubic_idx[ia];
            for( int q = 0; q < 3; ++q )
                g[3 * i + q] -= 2.0 * p->cubic_K[ia] * pow( sp[3 * i + q], 3.0 );
        }
}

/* Gradient_Exchange, :822-842 (OpenMP gather form) */
static void grad_exchange( orc_sys * s, const double * sp, double * g )
{
    const orc_params * p = &s->p;
#pragma omp parallel for
    for( int icell = 0; icell < s->n_cells_total; ++icell )
        for( int k = 0; k < p->n_exch; ++k )
        {
            int i = p->exch[k][0] + icell * p->n_cell_atoms;
            int j = orc_idx_from_pair( i, p->bc, p->n_cells, p->n_cell_atoms, p->exch[k] );
            if( j >= 0 )
                for( int q = 0; q < 3; ++q )
                    g[3 * i + q] -= p->exch_J[k] * sp[3 * j + q];
        }
}

/* Gradient_DMI, :844-864 (OpenMP gather form) */
static void grad_dmi( orc_sys * s, const double * sp, double * g )
{
    const orc_params * p = &s->p;
#pragma omp parallel for
    for( int icell = 0; icell < s->n_cells_total; ++icell )
        for( int k = 0; k < p->n_dmi; ++k )
        {
            int i = p->dmi[k][0] + icell * p->n_cell_atoms;
            int j = orc_idx_from_pair( i, p->bc, p->n_cells, p->n_cell_atoms, p->dmi[k] );
            if( j >= 0 )
            {
                double c[3];
                v3_cross( sp + 3 * j, p->dmi_normal[k], c );
                for( int q = 0; q < 3; ++q )
                    g[3 * i + q] -= p->dmi_D[k] * c[q];
            }
        }
}

/* the spins of quadruplet q anchored at cell icell (idx_from_pair from i's spin for j, k, l);
   returns 0 when one of them leaves an open boundary */
static int quad_spins( const orc_sys * s, int q, int icell, int out[4] )
{
    const orc_params * p = &s->p;
    const int * Q        = p->quad[q];
    int ispin            = Q[0] + icell * p->n_cell_atoms;
    int pj[5]            = { Q[0], Q[1], Q[4], Q[5], Q[6] };
    int pk[5]            = { Q[0], Q[2], Q[7], Q[8], Q[9] };
    int pl[5]            = { Q[0], Q[3], Q[10], Q[11], Q[12] };
    out[0]               = ispin;
    out[1]               = orc_idx_from_pair( ispin, p->bc, p->n_cells, p->n_cell_atoms, pj );
    out[2]               = orc_idx_from_pair( ispin, p->bc, p->n_cells, p->n_cell_atoms, pk );
    out[3]               = orc_idx_from_pair( ispin, p->bc, p->n_cells, p->n_cell_atoms, pl );
    return out[0] >= 0 && out[1] >= 0 && out[2] >= 0 && out[3] >= 0;
}

/* Gradient_Quadruplet, Hamiltonian_Heisenberg.cpp:1073-1117 (serial scatter over quadruplets, then
   cells; the cell loop runs a outer, c inner as da, db, dc there) */
static void grad_quadruplet( orc_sys * s, const double * sp, double * g )
{
    const orc_params * p = &s->p;
    for( int q = 0; q < p->n_quad; ++q )
        for( int da = 0; da < p->n_cells[0]; ++da )
            for( int db = 0; db < p->n_cells[1]; ++db )
                for( int dc = 0; dc < p->n_cells[2]; ++dc )
                {
                    int icell = da + p->n_cells[0] * ( db + p->n_cells[1] * dc );
                    int t[4];
                    if( !quad_spins( s, q, icell, t ) )
                        continue;
                    const double *si = sp + 3 * t[0], *sj = sp + 3 * t[1], *sk = sp + 3 * t[2], *sl = sp + 3 * t[3];
                    double Q = p->quad_Q[q], kl = v3_dot( sk, sl ), ij = v3_dot( si, sj );
                    for( int c = 0; c < 3; ++c )
                        g[3 * t[0] + c] -= Q * sj[c] * kl;
                    for( int c = 0; c < 3; ++c )
                        g[3 * t[1] + c] -= Q * si[c] * kl;
                    for( int c = 0; c < 3; ++c )
                        g[3 * t[2] + c] -= Q * ij * sl[c];
                    for( int c = 0; c < 3; ++c )
                        g[3 * t[3] + c] -= Q * ij * sk[c];
                }
}

void orc_energy_quadruplet( orc_sys * s, const double * sp, double * E )
{
    const orc_params * p = &s->p;
    for( int q = 0; q < p->n_quad; ++q )
        for( int da = 0; da < p->n_cells[0]; ++da )
            for( int db = 0; db < p->n_cells[1]; ++db )
                for( int dc = 0; dc < p->n_cells[2]; ++dc )
                {
                    int icell = da + p->n_cells[0] * ( db + p->n_cells[1] * dc );
                    int t[4];
                    if( !quad_spins( s, q, icell, t ) )
                        continue;
                    double e = 0.25 * p->quad_Q[q] * v3_dot( sp + 3 * t[0], sp + 3 * t[1] )
                               * v3_dot( sp + 3 * t[2], sp + 3 * t[3] );
                    for( int r = 0; r < 4; ++r )
                        E[t[r]] -= e;
                }
}

static void grad_ddi( orc_sys * s, const double * sp, double * g )
{
    if( s->p.ddi_method == ORC_DDI_FFT )
        gradient_ddi_fft( s, sp, g );
    else if( s->p.ddi_method == ORC_DDI_CUTOFF && s->p.ddi_radius < 0 )
        gradient_ddi_direct( s, sp, g );
    else if( s->p.ddi_method == ORC_DDI_CUTOFF )
        gradient_ddi_cutoff( s, sp, g );
}

void orc_gradient_term( orc_sys * s, int term, const double * spins, double * grad )
{
    switch( term )
    {
        case 0: grad_zeeman( s, grad ); break;
        case 1: grad_aniso( s, spins, grad ); break;
        case 2: grad_cubic( s, spins, grad ); break;
        case 3: grad_exchange( s, spins, grad ); break;
        case 4: grad_dmi( s, spins, grad ); break;
        case 5: grad_ddi( s, spins, grad ); break;
        case 6: gradient_ddi_direct( s, spins, grad ); break;
        case 7: grad_quadruplet( s, spins, grad ); break;
        default: break;
    }
}

static int has_zeeman( const orc_sys * s )
{
    return fabs( s->p.field_tesla * ORC_MU_B ) > 1e-60;
}

/* Hamiltonian_Heisenberg.cpp:670-702 */
void orc_gradient( orc_sys * s, const double * spins, double * grad )
{
    memset( grad, 0, sizeof( double ) * 3 * (size_t)s->nos );
    if( has_zeeman( s ) )
        grad_zeeman( s, grad );
    if( s->p.n_aniso > 0 || s->site_K )
        grad_aniso( s, spins, grad );
    if( s->p.n_cubic > 0 )
        grad_cubic( s, spins, grad );
    if( s->p.n_exch > 0 )
        grad_exchange( s, spins, grad );
    if( s->p.n_dmi > 0 )
        grad_dmi( s, spins, grad );
    if( s->p.ddi_method != ORC_DDI_NONE )
        grad_ddi( s, spins, grad );
    if( s->p.n_quad > 0 )
        grad_quadruplet( s, spins, grad );
}

/* Deterministic parallel sum: fixed blocks of ORC_SUM_BLOCK terms summed in parallel, the
 * block sums added in index order, so the result does not depend on the thread count or on
 * the order in which threads finish (an OpenMP reduction clause guarantees neither). */
#define ORC_SUM_BLOCK 4096
struct sum_ctx
{
    const orc_sys * s;
    const double * spins;
    const double * grad;
    const double * ext;
};
static double det_sum( long n, double ( *term )( const struct sum_ctx *, long ), const struct sum_ctx * ctx )
{
    const long nb  = ( n + ORC_SUM_BLOCK - 1 ) / ORC_SUM_BLOCK;
    double * part  = (double *)malloc( sizeof( double ) * (size_t)( nb > 0 ? nb : 1 ) );
#pragma omp parallel for schedule( static )
    for( long b = 0; b < nb; ++b )
    {
        const long i1 = ( b + 1 ) * ORC_SUM_BLOCK < n ? ( b + 1 ) * ORC_SUM_BLOCK : n;
        double t      = 0;
        for( long i = b * ORC_SUM_BLOCK; i < i1; ++i )
            t += term( ctx, i );
        part[b] = t;
    }
    double total = 0;
    for( long b = 0; b < nb; ++b )
        total += part[b];
    free( part );
    return total;
}
/* 1/2 g_i . s_i (Hamiltonian_Heisenberg.cpp:732) */
static double term_half_g_dot_s( const struct sum_ctx * c, long i )
{
    return 0.5 * v3_dot( c->grad + 3 * i, c->spins + 3 * i );
}
/* cubic anisotropy energy of one cell, -K/2 (sx^4 + sy^4 + sz^4) per cubic atom (:734-742) */
static double term_cubic_cell( const struct sum_ctx * c, long icell )
{
    const orc_params * p = &c->s->p;
    double e             = 0;
    for( int ia = 0; ia < p->n_cubic; ++ia )
    {
        long i           = icell * p->n_cell_atoms + p->cubic_idx[ia];
        const double * v = c->spins + 3 * i;
        e -= p->cubic_K[ia] / 2 * ( pow( v[0], 4.0 ) + pow( v[1], 4.0 ) + pow( v[2], 4.0 ) );
    }
    return e;
}
/* Zeeman energy -mu_i B . s_i (:749-750) */
static double term_zeeman( const struct sum_ctx * c, long i )
{
    return -c->s->mu_s[i] * v3_dot( c->ext, c->spins + 3 * i );
}

/* Hamiltonian_Heisenberg.cpp:704-766 */
void orc_gradient_and_energy( orc_sys * s, const double * spins, double * grad, double * energy )
{
    memset( grad, 0, sizeof( double ) * 3 * (size_t)s->nos );
    double E = 0;
    if( s->p.n_aniso > 0 || s->site_K )
        grad_aniso( s, spins, grad );
    if( s->p.n_exch > 0 )
        grad_exchange( s, spins, grad );
    if( s->p.n_dmi > 0 )
        grad_dmi( s, spins, grad );
    if( s->p.ddi_method != ORC_DDI_NONE )
        grad_ddi( s, spins, grad );
    E += det_sum( s->nos, term_half_g_dot_s, &( struct sum_ctx ){ s, spins, grad, NULL } );
    if( s->p.n_cubic > 0 )
    {
        grad_cubic( s, spins, grad );
        E += det_sum( s->n_cells_total, term_cubic_cell, &( struct sum_ctx ){ s, spins, grad, NULL } );
    }
    if( has_zeeman( s ) )
    {
        double B = s->p.field_tesla * ORC_MU_B;
        double ext[3];
        for( int q = 0; q < 3; ++q )
            ext[q] = s->p.field_normal[q] * B;
        grad_zeeman( s, grad );
        E += det_sum( s->nos, term_zeeman, &( struct sum_ctx ){ s, spins, grad, ext } );
    }
    /* quadruplets: the reference tests idx_quadruplet > 0 (:754), so they count only when another
       contribution precedes them in the list of Update_Energy_Contributions (:200-260) */
    const int others = has_zeeman( s ) || s->p.n_aniso > 0 || s->site_K || s->p.n_cubic > 0 || s->p.n_exch > 0
                       || s->p.n_dmi > 0 || s->p.ddi_method != ORC_DDI_NONE;
    if( s->p.n_quad > 0 && others )
    {
        grad_quadruplet( s, spins, grad );
        double * eq = (double *)calloc( (size_t)s->nos, sizeof( double ) );
        orc_energy_quadruplet( s, spins, eq );
        double sum = 0;
        for( int i = 0; i < s->nos; ++i )
            sum += eq[i];
        free( eq );
        E += sum;
    }
    *energy = E;
}

/* ------------------------------------------------------------------ */
/* RNG: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller              */
/* ------------------------------------------------------------------ */
static void philox4x32_10( uint32_t ctr[4], uint32_t key[2] )
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t k0 = key[0], k1 = key[1];
    for( int r = 0; r < 10; ++r )
    {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)( p0 >> 32 ), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)( p1 >> 32 ), lo1 = (uint32_t)p1;
        uint32_t n0  = hi1 ^ ctr[1] ^ k0;
        uint32_t n1  = lo1;
        uint32_t n2  = hi0 ^ ctr[3] ^ k1;
        uint32_t n3  = lo0;
        ctr[0]       = n0;
        ctr[1]       = n1;
        ctr[2]       = n2;
        ctr[3]       = n3;
        k0 += W0;
        k1 += W1;
    }
}

static inline double u53( uint32_t a, uint32_t b )
{
    return ( (double)( a >> 5 ) * 67108864.0 + (double)( b >> 6 ) ) * ( 1.0 / 9007199254740992.0 );
}

void orc_normal3( uint64_t seed, uint64_t step, uint64_t site, double out[3] )
{
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)( seed >> 32 ) };
    uint32_t c0[4]  = { (uint32_t)site, (uint32_t)( site >> 32 ), (uint32_t)step, (uint32_t)( step >> 32 ) };
    uint32_t c1[4]  = { c0[0], c0[1] | 0x80000000u, c0[2], c0[3] };
    philox4x32_10( c0, key );
    philox4x32_10( c1, key );
    double u1 = 1.0 - u53( c0[0], c0[1] );
    double u2 = u53( c0[2], c0[3] );
    double u3 = 1.0 - u53( c1[0], c1[1] );
    double u4 = u53( c1[2], c1[3] );
    double r1 = sqrt( -2.0 * log( u1 ) );
    double r2 = sqrt( -2.0 * log( u3 ) );
    out[0]    = r1 * cos( 2.0 * ORC_PI * u2 );
    out[1]    = r1 * sin( 2.0 * ORC_PI * u2 );
    out[2]    = r2 * cos( 2.0 * ORC_PI * u4 );
}

/* ------------------------------------------------------------------ */
/* LLG                                                                 */
/* ------------------------------------------------------------------ */
void orc_rotate( const double v[3], const double axis[3], double angle, double out[3] )
{
    double c = cos( angle ), sn = sin( angle );
    double cr[3];
    v3_cross( axis, v, cr );
    double d = v3_dot( axis, v );
    double o[3];
    for( int q = 0; q < 3; ++q )
        o[q] = v[q] * c + cr[q] * sn + axis[q] * d * ( 1 - c );
    out[0] = o[0], out[1] = o[1], out[2] = o[2];
}

void orc_sib_transform( const double s[3], const double f[3], double out[3] )
{
    double e1[3] = { s[0], s[1], s[2] };
    double A[3]  = { 0.5 * f[0], 0.5 * f[1], 0.5 * f[2] };
    double detAi = 1.0 / ( 1 + pow( sqrt( v3_sqnorm( A ) ), 2.0 ) );
    double c[3], a2[3];
    v3_cross( e1, A, c );
    for( int q = 0; q < 3; ++q )
        a2[q] = e1[q] - c[q];
    double o0 = ( a2[0] * ( A[0] * A[0] + 1 ) + a2[1] * ( A[0] * A[1] - A[2] ) + a2[2] * ( A[0] * A[2] + A[1] ) ) * detAi;
    double o1 = ( a2[0] * ( A[1] * A[0] + A[2] ) + a2[1] * ( A[1] * A[1] + 1 ) + a2[2] * ( A[1] * A[2] - A[0] ) ) * detAi;
    double o2 = ( a2[0] * ( A[2] * A[0] - A[1] ) + a2[1] * ( A[2] * A[1] + A[0] ) + a2[2] * ( A[2] * A[2] + 1 ) ) * detAi;
    out[0] = o0, out[1] = o1, out[2] = o2;
}

/* Geometry::calculateBounds, core/src/data/Geometry.cpp:729-743: the bounds start at 0 */
static void geometry_bounds( const orc_sys * s, double bmin[3], double bmax[3] )
{
    for( int q = 0; q < 3; ++q )
        bmin[q] = bmax[q] = 0;
    for( int i = 0; i < s->nos; ++i )
        for( int q = 0; q < 3; ++q )
        {
            double v = s->positions[3 * i + q];
            if( v < bmin[q] )
                bmin[q] = v;
            if( v > bmax[q] )
                bmax[q] = v;
        }
}

/* Vectormath::get_gradient_distribution, Vectormath.cpp:633-652 */
void orc_temperature_distribution( const orc_sys * s, const double direction[3], double T, double inclination,
                                   double * out )
{
    double d[3] = { direction[0], direction[1], direction[2] };
    double n    = sqrt( d[0] * d[0] + d[1] * d[1] + d[2] * d[2] );
    if( n > 0 )
        for( int q = 0; q < 3; ++q )
            d[q] /= n; /* :638 */
    double bmin[3], bmax[3];
    geometry_bounds( s, bmin, bmax );
    /* :641 distribution_i = inclination * (d . r_i) */
    for( int i = 0; i < s->nos; ++i )
    {
        const double * r = s->positions + 3 * i;
        out[i]           = inclination * ( d[0] * r[0] + d[1] * r[1] + d[2] * r[2] );
    }
    /* :644-648 the start: T at the smaller of the bounds' projections */
    double pmin     = bmin[0] * d[0] + bmin[1] * d[1] + bmin[2] * d[2];
    double pmax     = bmax[0] * d[0] + bmax[1] * d[1] + bmax[2] * d[2];
    double dist_min = pmin < pmax ? pmin : pmax;
    double shift    = T - inclination * dist_min;
    for( int i = 0; i < s->nos; ++i )
    {
        double t = out[i] + shift;
        /* :651 set_range(0, 1e30) */
        out[i] = t < 0 ? 0 : ( t > 1e30 ? 1e30 : t );
    }
}

/* Vectormath::jacobian, Vectormath.cpp:816-898, then s_c_grad = J je (Method_LLG.cpp:188-196) */
void orc_jacobian_direction( const orc_sys * s, const double * vf, const double je[3], double * out )
{
    const orc_params * p = &s->p;
    const int N = p->n_cell_atoms, n[3] = { p->n_cells[0], p->n_cells[1], p->n_cells[2] };
    /* :827-838 base matrix of the scaled Bravais vectors (as columns) and its inverse */
    double B[3][3], Bi[3][3];
    for( int r = 0; r < 3; ++r )
        for( int k = 0; k < 3; ++k )
            B[r][k] = p->lattice_constant * p->bravais[k][r];
    double det = B[0][0] * ( B[1][1] * B[2][2] - B[1][2] * B[2][1] ) - B[0][1] * ( B[1][0] * B[2][2] - B[1][2] * B[2][0] )
                 + B[0][2] * ( B[1][0] * B[2][1] - B[1][1] * B[2][0] );
    Bi[0][0] = ( B[1][1] * B[2][2] - B[1][2] * B[2][1] ) / det;
    Bi[0][1] = ( B[0][2] * B[2][1] - B[0][1] * B[2][2] ) / det;
    Bi[0][2] = ( B[0][1] * B[1][2] - B[0][2] * B[1][1] ) / det;
    Bi[1][0] = ( B[1][2] * B[2][0] - B[1][0] * B[2][2] ) / det;
    Bi[1][1] = ( B[0][0] * B[2][2] - B[0][2] * B[2][0] ) / det;
    Bi[1][2] = ( B[0][2] * B[1][0] - B[0][0] * B[1][2] ) / det;
    Bi[2][0] = ( B[1][0] * B[2][1] - B[1][1] * B[2][0] ) / det;
    Bi[2][1] = ( B[0][1] * B[2][0] - B[0][0] * B[2][1] ) / det;
    Bi[2][2] = ( B[0][0] * B[1][1] - B[0][1] * B[1][0] ) / det;
#pragma omp parallel for collapse( 3 )
    for( int c = 0; c < n[2]; ++c )
        for( int b = 0; b < n[1]; ++b )
            for( int a = 0; a < n[0]; ++a )
                for( int ib = 0; ib < N; ++ib )
                {
                    const int pos[3] = { a, b, c };
                    const long cur   = ib + (long)N * ( a + (long)n[0] * ( b + (long)n[1] * c ) );
                    double Mcol[3][3]; /* Mcol[t] = column t of m_matrix */
                    for( int t = 0; t < 3; ++t )
                    {
                        /* linear_idx (Vectormath.cpp:773-814) at +-1 along translation t */
                        long idx[2];
                        for( int side = 0; side < 2; ++side )
                        {
                            int q[3] = { pos[0], pos[1], pos[2] };
                            q[t] += side == 0 ? 1 : -1;
                            if( q[t] < 0 || q[t] >= n[t] )
                            {
                                if( p->bc[t] )
                                    q[t] = ( n[t] + ( q[t] % n[t] ) ) % n[t];
                                else
                                {
                                    idx[side] = -1;
                                    continue;
                                }
                            }
                            idx[side] = ib + (long)N * ( q[0] + (long)n[0] * ( q[1] + (long)n[1] * q[2] ) );
                        }
                        double factor = 0.5; /* central difference */
                        const double *m0, *m1;
                        if( idx[0] >= 0 )
                            m0 = vf + 3 * idx[0];
                        else
                        {
                            m0 = vf + 3 * cur;
                            factor *= 2;
                        }
                        if( idx[1] >= 0 )
                            m1 = vf + 3 * idx[1];
                        else
                        {
                            m1 = vf + 3 * cur;
                            factor *= 2;
                        }
                        for( int r = 0; r < 3; ++r )
                            Mcol[t][r] = factor * ( m0[r] - m1[r] );
                    }
                    /* :896 J = m_matrix * B^-1, then J je */
                    for( int r = 0; r < 3; ++r )
                    {
                        double v = 0;
                        for( int k = 0; k < 3; ++k )
                        {
                            double Jrk = Mcol[0][r] * Bi[0][k] + Mcol[1][r] * Bi[1][k] + Mcol[2][r] * Bi[2][k];
                            v += Jrk * je[k];
                        }
                        out[3 * cur + r] = v;
                    }
                }
}

/* Method_LLG::Prepare_Thermal_Field, Method_LLG.cpp:66-110 (generator replaced, see header) */
static void prepare_thermal( orc_sys * s, const orc_llg * par, uint64_t step, int local_step )
{
    if( !( par->temperature > 0 || par->tgrad_inclination != 0 ) )
        return;
    double a   = par->damping;
    double eps = sqrt( 2 * a * par->dt * ORC_GAMMA / ORC_MU_B * ORC_K_B ) / ( 1 + a * a );
    /* :80-93: with a gradient, the site's temperature from the distribution */
    double * Ti = NULL;
    if( par->tgrad_inclination != 0 )
    {
        Ti = (double *)malloc( sizeof( double ) * (size_t)s->nos );
        orc_temperature_distribution( s, par->tgrad_direction, par->temperature, par->tgrad_inclination, Ti );
    }
#pragma omp parallel for
    for( int i = 0; i < s->nos; ++i )
    {
        double z[3];
        if( par->noise )
        {
            const double * src = par->noise + ( (size_t)local_step * s->nos + i ) * 3;
            z[0] = src[0], z[1] = src[1], z[2] = src[2];
        }
        else
            orc_normal3( par->seed, step, (uint64_t)i, z );
        double f = eps * sqrt( ( Ti ? Ti[i] : par->temperature ) / s->mu_s[i] );
        for( int q = 0; q < 3; ++q )
            s->xi[3 * i + q] = f * z[q];
    }
    free( Ti );
}

/* Method_LLG::Calculate_Force, Method_LLG.cpp:113-129 */
static double calc_force( orc_sys * s, const double * spins, double * force )
{
    double E;
    orc_gradient_and_energy( s, spins, s->g, &E );
    for( long t = 0; t < 3L * s->nos; ++t )
        force[t] = -1 * s->g[t];
    return E;
}

/* Method_LLG::Calculate_Force_Virtual, Method_LLG.cpp:132-226 */
static void calc_force_virtual( orc_sys * s, const orc_llg * par, int solver, const double * spins, const double * F,
                                double * Fv )
{
    double a = par->damping;
    if( solver == ORC_SOLVER_LBFGS_OSO || solver == ORC_SOLVER_LBFGS_ATLAS || par->direct_minimization
        || solver == ORC_SOLVER_VP || solver == ORC_SOLVER_VP_OSO )
    {
        /* Method_LLG.cpp:160-171: s x F, scaled by dt gamma / mu_B except for the L-BFGS solvers */
        const int lbfgs = solver == ORC_SOLVER_LBFGS_OSO || solver == ORC_SOLVER_LBFGS_ATLAS;
        double dtg      = lbfgs ? 1.0 : par->dt * ORC_GAMMA / ORC_MU_B;
#pragma omp parallel for
        for( int i = 0; i < s->nos; ++i )
        {
            double c[3];
            v3_cross( spins + 3 * i, F + 3 * i, c );
            for( int q = 0; q < 3; ++q )
                Fv[3 * i + q] = dtg * c[q];
        }
        return;
    }
    double dtg = par->dt * ORC_GAMMA / ORC_MU_B / ( 1 + a * a );
    int thermal = par->temperature > 0 || par->tgrad_inclination != 0;
    /* STT (:150-157, 182-212): a_j = b_j = stt_magnitude, the current direction je = s_c_vec */
    const double a_j = par->stt_magnitude, beta = par->beta;
    const int stt    = a_j > 0;
    double * s_c_grad = NULL;
    if( stt && par->stt_use_gradient )
    {
        s_c_grad = (double *)malloc( sizeof( double ) * 3 * (size_t)s->nos );
        orc_jacobian_direction( s, spins, par->stt_normal, s_c_grad );
    }
#pragma omp parallel for
    for( int i = 0; i < s->nos; ++i )
    {
        double c[3], v[3];
        for( int q = 0; q < 3; ++q )
            v[q] = dtg * F[3 * i + q];
        v3_cross( spins + 3 * i, F + 3 * i, c );
        for( int q = 0; q < 3; ++q )
            v[q] += dtg * a * c[q];
        for( int q = 0; q < 3; ++q )
            v[q] /= s->mu_s[i];
        if( stt && par->stt_use_gradient )
        {
            /* :197-201: + dtg a_j (alpha - beta) g + dtg a_j (1 + beta alpha) g x s */
            const double * g = s_c_grad + 3 * i;
            const double c1 = dtg * a_j * ( a - beta ), c2 = dtg * a_j * ( 1 + beta * a );
            for( int q = 0; q < 3; ++q )
                v[q] += c1 * g[q];
            v3_cross( g, spins + 3 * i, c );
            for( int q = 0; q < 3; ++q )
                v[q] += c2 * c[q];
        }
        else if( stt )
        {
            /* :206-207: - dtg a_j (alpha - beta) s_c - dtg a_j (1 + beta alpha) s_c x s */
            const double c1 = -dtg * a_j * ( a - beta ), c2 = -dtg * a_j * ( 1 + beta * a );
            for( int q = 0; q < 3; ++q )
                v[q] += c1 * par->stt_normal[q];
            v3_cross( par->stt_normal, spins + 3 * i, c );
            for( int q = 0; q < 3; ++q )
                v[q] += c2 * c[q];
        }
        if( thermal )
        {
            const double * xi = s->xi + 3 * i;
            for( int q = 0; q < 3; ++q )
                v[q] += 1 * xi[q];
            v3_cross( spins + 3 * i, xi, c );
            for( int q = 0; q < 3; ++q )
                v[q] += a * c[q];
        }
        for( int q = 0; q < 3; ++q )
            Fv[3 * i + q] = v[q];
    }
    free( s_c_grad );
}

/* -(a x b) helper: out = -1 * a.cross(b) (set_c_cross(-1, a, b, out)) */
static inline void neg_cross( const double * a, const double * b, double * o )
{
    double c[3];
    v3_cross( a, b, c );
    o[0] = -1 * c[0];
    o[1] = -1 * c[1];
    o[2] = -1 * c[2];
}

/* ------------------------------------------------------------------ */
/* Direct-minimisation solvers: VP (Solver_VP.hpp:29-114), VP_OSO      */
/* (Solver_VP_OSO.hpp:33-110), LBFGS_OSO (Solver_LBFGS_OSO.hpp:37-72),  */
/* LBFGS_Atlas (Solver_LBFGS_Atlas.hpp:44-101) with the kernels of     */
/* Solver_Kernels.cpp:50-244 and Solver_Kernels.hpp:44-191.            */
/* ------------------------------------------------------------------ */
#define ORC_LBFGS_MEMORY 3 /* n_lbfgs_memory, Solver_LBFGS_OSO.hpp:15 */

typedef struct min_state
{
    int dim;             /* 3: cartesian / OSO tangent vectors; 2: atlas coordinates */
    double * v;          /* VP, VP_OSO: velocities */
    double * prev;       /* VP: previous force; VP_OSO: previous gradient */
    double * grad;       /* VP_OSO, L-BFGS: gradient (OSO) or atlas residual */
    double * grad_pr;    /* L-BFGS: previous gradient */
    double * sd;         /* search direction */
    double * q;          /* L-BFGS two-loop vector */
    double * da[ORC_LBFGS_MEMORY], * dg[ORC_LBFGS_MEMORY];
    double rho[ORC_LBFGS_MEMORY], alpha[ORC_LBFGS_MEMORY];
    int local_iter;
    double maxmove;
    double * a3;         /* LBFGS_Atlas: chart (+1 / -1) of each spin */
} min_state;

static int is_minimiser( int solver )
{
    return solver == ORC_SOLVER_VP || solver == ORC_SOLVER_VP_OSO || solver == ORC_SOLVER_LBFGS_OSO
           || solver == ORC_SOLVER_LBFGS_ATLAS;
}

/* Method_Solver<...>::Initialize of each solver (fresh state, as a new method object) */
static min_state * min_state_create( int solver, int nos, const double * spins )
{
    min_state * st = (min_state *)calloc( 1, sizeof( min_state ) );
    st->dim        = solver == ORC_SOLVER_LBFGS_ATLAS ? 2 : 3;
    size_t n       = (size_t)st->dim * (size_t)nos;
    st->v          = (double *)calloc( n, sizeof( double ) );
    st->prev       = (double *)calloc( n, sizeof( double ) );
    st->grad       = (double *)calloc( n, sizeof( double ) );
    st->grad_pr    = (double *)calloc( n, sizeof( double ) );
    st->sd         = (double *)calloc( n, sizeof( double ) );
    st->q          = (double *)calloc( n, sizeof( double ) );
    for( int k = 0; k < ORC_LBFGS_MEMORY; ++k )
    {
        st->da[k] = (double *)calloc( n, sizeof( double ) );
        st->dg[k] = (double *)calloc( n, sizeof( double ) );
    }
    st->maxmove = solver == ORC_SOLVER_LBFGS_ATLAS ? 0.05 : ORC_PI / 200.0; /* Solver_LBFGS_*.hpp */
    if( solver == ORC_SOLVER_LBFGS_ATLAS )
    {
        /* Solver_LBFGS_Atlas.hpp:34-41 */
        st->a3 = (double *)calloc( (size_t)nos, sizeof( double ) );
        for( int i = 0; i < nos; ++i )
            st->a3[i] = spins[3 * i + 2] > 0 ? 1.0 : -1.0;
    }
    return st;
}

static void min_state_destroy( min_state * st )
{
    if( !st )
        return;
    free( st->v );
    free( st->prev );
    free( st->grad );
    free( st->grad_pr );
    free( st->sd );
    free( st->q );
    for( int k = 0; k < ORC_LBFGS_MEMORY; ++k )
    {
        free( st->da[k] );
        free( st->dg[k] );
    }
    free( st->a3 );
    free( st );
}

/* Vectormath::dot over a vector field (per-spin dot products, summed deterministically) */
static double term_dot3( const struct sum_ctx * c, long i )
{
    return v3_dot( c->spins + 3 * i, c->grad + 3 * i );
}
static double term_dot2( const struct sum_ctx * c, long i )
{
    const double *a = c->spins + 2 * i, *b = c->grad + 2 * i;
    return a[0] * b[0] + a[1] * b[1];
}
static double field_dot( const double * a, const double * b, int nos, int dim )
{
    struct sum_ctx c = { NULL, a, b, NULL };
    return det_sum( nos, dim == 3 ? term_dot3 : term_dot2, &c );
}

/* Solver_Kernels::lbfgs_get_searchdir, Solver_Kernels.hpp:51-191 (one image) */
static void lbfgs_searchdir( min_state * st, int nos )
{
    const int M      = ORC_LBFGS_MEMORY;
    const long n     = (long)st->dim * nos;
    const double eps = 1e-300;
    const int m      = st->local_iter % M;
    if( st->local_iter == 0 )
    {
        for( long j = 0; j < n; ++j )
        {
            st->grad_pr[j] = st->grad[j];
            st->sd[j]      = -st->grad[j];
        }
        for( int k = 0; k < M; ++k )
        {
            st->rho[k] = 0.0;
            memset( st->da[k], 0, sizeof( double ) * (size_t)n );
            memset( st->dg[k], 0, sizeof( double ) * (size_t)n );
        }
    }
    else
    {
        for( long j = 0; j < n; ++j )
        {
            st->da[m][j] = st->sd[j];
            st->dg[m][j] = st->grad[j] - st->grad_pr[j];
        }
        double rinv = field_dot( st->dg[m], st->da[m], nos, st->dim );
        if( rinv > eps )
            st->rho[m] = 1.0 / rinv;
        else
        {
            st->local_iter = 0;
            lbfgs_searchdir( st, nos );
            return;
        }
        memcpy( st->q, st->grad, sizeof( double ) * (size_t)n );
        for( int k = M - 1; k > -1; --k )
        {
            const int c  = ( k + m + 1 ) % M;
            double temp  = field_dot( st->da[c], st->q, nos, st->dim );
            st->alpha[c] = st->rho[c] * temp;
            for( long j = 0; j < n; ++j )
                st->q[j] += -st->alpha[c] * st->dg[c][j];
        }
        double dy2    = field_dot( st->dg[m], st->dg[m], nos, st->dim );
        double rhody2 = dy2 * st->rho[m];
        double inv    = rhody2 > eps ? 1.0 / rhody2 : 1.0 / ( eps );
        for( long j = 0; j < n; ++j )
            st->sd[j] = inv * st->q[j];
        for( int k = 0; k < M; ++k )
        {
            const int c   = st->local_iter < M ? k : ( k + m + 1 ) % M;
            double rhopdg = field_dot( st->dg[c], st->sd, nos, st->dim );
            rhopdg *= st->rho[c];
            for( long j = 0; j < n; ++j )
                st->sd[j] += ( st->alpha[c] - rhopdg ) * st->da[c][j];
        }
        for( long j = 0; j < n; ++j )
        {
            st->grad_pr[j] = st->grad[j];
            st->sd[j]      = -st->sd[j];
        }
    }
    st->local_iter++;
}

/* Step scaling to the maximum move from the largest per-image rms of the search direction:
   the min over images of maximum_rotation (Solver_LBFGS_OSO.hpp:59-62, Solver_Kernels.cpp:95-103)
   and the max-rms rule of Solver_LBFGS_Atlas.hpp:74-86 give the same factor. */
static double rms_scaling( const min_state * st, int noi, int nos )
{
    double rms_max = 0;
    for( int img = 0; img < noi; ++img )
    {
        const double * d = st->sd + (size_t)img * st->dim * nos;
        double rms       = sqrt( field_dot( d, d, nos, st->dim ) / nos );
        if( rms > rms_max )
            rms_max = rms;
    }
    return rms_max > st->maxmove ? st->maxmove / rms_max : 1.0;
}

/* Solver_Kernels::oso_rotate, Solver_Kernels.cpp:63-93 (one spin) */
static void oso_rotate_spin( double * sp, const double * sd )
{
    double theta = sqrt( v3_sqnorm( sd ) );
    double q = cos( theta ), w = 1 - q, x = -sd[0] / theta, y = -sd[1] / theta, z = -sd[2] / theta,
           s1 = -y * z * w, s2 = x * z * w, s3 = -x * y * w, p1 = x * sin( theta ), p2 = y * sin( theta ),
           p3 = z * sin( theta );
    if( theta > 1.0e-20 )
    {
        double t1 = ( q + z * z * w ) * sp[0] + ( s1 + p1 ) * sp[1] + ( s2 + p2 ) * sp[2];
        double t2 = ( s1 - p1 ) * sp[0] + ( q + y * y * w ) * sp[1] + ( s3 + p3 ) * sp[2];
        double t3 = ( s2 - p2 ) * sp[0] + ( s3 - p3 ) * sp[1] + ( q + x * x * w ) * sp[2];
        sp[0]     = t1;
        sp[1]     = t2;
        sp[2]     = t3;
    }
}

/* the OSO gradient t (-s x F), t = ((0,0,1),(0,-1,0),(1,0,0)) (Solver_Kernels.cpp:50-61) */
static void oso_gradient( const double * sp, const double * F, double * g )
{
    double c[3];
    v3_cross( sp, F, c );
    const double w[3] = { -c[0], -c[1], -c[2] };
    g[0] = w[2];
    g[1] = -w[1];
    g[2] = w[0];
}

/* per-spin sums of the atlas chart change (lbfgs_atlas_transform_direction's rho update) */
struct atlas_ctx
{
    const double * spins;
    const double * a3;
    const double * da;
    const double * dg;
};
static double term_atlas_rho( const struct sum_ctx * c, long i )
{
    const struct atlas_ctx * a = (const struct atlas_ctx *)c->ext;
    const double sz            = a->spins[3 * i + 2];
    if( !( sz * a->a3[i] < 0 ) )
        return 0.0;
    const double a3n    = sz > 0 ? 1 : -1;
    const double factor = ( 1 - a3n * sz ) / ( 1 + a3n * sz );
    return ( factor * factor - 1 ) * ( a->da[2 * i] * a->dg[2 * i] + a->da[2 * i + 1] * a->dg[2 * i + 1] );
}

/* One update of a direct-minimisation solver on noi images of nos spins (contiguous AoS
   spins), given their forces F (the LLG force -g, or GNEB's F_total).  dt: the images' LLG
   time step.  gtmp: 3 noi nos doubles of scratch. */
static void min_update( int solver, double dt, int noi, int nos, double * spins, const double * F, min_state * st,
                        double * gtmp )
{
    const int ns      = noi * nos; /* spins of all images */
    const double mass = 1.0;       /* Method_Solver.hpp:174 */
    if( solver == ORC_SOLVER_VP || solver == ORC_SOLVER_VP_OSO )
    {
        /* VP on the force (VP) or on the OSO gradient's negative (VP_OSO) */
        double * g = solver == ORC_SOLVER_VP ? gtmp : st->grad;
        if( solver == ORC_SOLVER_VP )
            memcpy( g, F, sizeof( double ) * 3 * (size_t)ns );
        else
        {
            for( int i = 0; i < ns; ++i )
            {
                oso_gradient( spins + 3 * i, F + 3 * i, g + 3 * i );
                for( int q = 0; q < 3; ++q )
                    g[3 * i + q] *= -1.0;
            }
        }
        for( long j = 0; j < 3L * ns; ++j )
            st->v[j] += 0.5 / mass * ( st->prev[j] + g[j] );
        double proj  = field_dot( st->v, g, ns, 3 );
        double fn2   = field_dot( g, g, ns, 3 );
        double ratio = proj / fn2;
        for( long j = 0; j < 3L * ns; ++j )
            st->v[j] = proj <= 0 ? 0.0 : g[j] * ratio;
        if( solver == ORC_SOLVER_VP )
        {
            for( int i = 0; i < ns; ++i )
            {
                double t[3];
                for( int q = 0; q < 3; ++q )
                    t[q] = spins[3 * i + q] + dt * st->v[3 * i + q] + 0.5 / mass * dt * g[3 * i + q];
                v3_normalize( t );
                for( int q = 0; q < 3; ++q )
                    spins[3 * i + q] = t[q];
            }
        }
        else
        {
            for( int i = 0; i < ns; ++i )
            {
                for( int q = 0; q < 3; ++q )
                    st->sd[3 * i + q] = dt * st->v[3 * i + q] + 0.5 / mass * dt * g[3 * i + q];
                oso_rotate_spin( spins + 3 * i, st->sd + 3 * i );
            }
        }
        memcpy( st->prev, g, sizeof( double ) * 3 * (size_t)ns );
        return;
    }
    if( solver == ORC_SOLVER_LBFGS_OSO )
    {
        for( int i = 0; i < ns; ++i )
            oso_gradient( spins + 3 * i, F + 3 * i, st->grad + 3 * i );
        lbfgs_searchdir( st, ns );
        double scaling = rms_scaling( st, noi, nos );
        for( int i = 0; i < ns; ++i )
        {
            for( int q = 0; q < 3; ++q )
                st->sd[3 * i + q] *= scaling;
            oso_rotate_spin( spins + 3 * i, st->sd + 3 * i );
        }
        return;
    }
    /* LBFGS_Atlas */
    for( int i = 0; i < ns; ++i )
    {
        /* Solver_Kernels::atlas_calc_gradients, Solver_Kernels.cpp:129-152 */
        const double * sp = spins + 3 * i;
        const double * f  = F + 3 * i;
        const double a3   = st->a3[i];
        double J00 = sp[1] * sp[1] + sp[2] * ( sp[2] + a3 );
        double J10 = -sp[0] * sp[1];
        double J01 = -sp[0] * sp[1];
        double J11 = sp[0] * sp[0] + sp[2] * ( sp[2] + a3 );
        double J02 = -sp[0] * ( sp[2] + a3 );
        double J12 = -sp[1] * ( sp[2] + a3 );
        st->grad[2 * i]     = -( J00 * f[0] + J01 * f[1] + J02 * f[2] );
        st->grad[2 * i + 1] = -( J10 * f[0] + J11 * f[1] + J12 * f[2] );
    }
    lbfgs_searchdir( st, ns );
    double scaling = rms_scaling( st, noi, nos );
    for( int i = 0; i < ns; ++i )
    {
        double * d = st->sd + 2 * i;
        d[0] *= scaling;
        d[1] *= scaling;
        /* Solver_Kernels::atlas_rotate, Solver_Kernels.cpp:105-127 */
        double * sp        = spins + 3 * i;
        const double a3    = st->a3[i];
        const double gamma = ( 1 + sp[2] * a3 );
        const double denom = ( sp[0] * sp[0] + sp[1] * sp[1] ) / gamma + 2 * ( d[0] * sp[0] + d[1] * sp[1] )
                             + gamma * ( d[0] * d[0] + d[1] * d[1] );
        sp[0]          = 2 * ( sp[0] + d[0] * gamma );
        sp[1]          = 2 * ( sp[1] + d[1] * gamma );
        sp[2]          = a3 * ( gamma - denom );
        const double r = 1 / ( gamma + denom );
        sp[0] *= r;
        sp[1] *= r;
        sp[2] *= r;
    }
    /* ncg_atlas_check_coordinates, Solver_Kernels.cpp:154-180 (tol -0.6); the reference
       compares the spins of image 0 with every image's charts */
    int any_flip = 0;
    for( int img = 0; img < noi && !any_flip; ++img )
        for( int i = 0; i < nos; ++i )
            if( spins[3 * i + 2] * st->a3[(size_t)img * nos + i] < -0.6 )
            {
                any_flip = 1;
                break;
            }
    if( any_flip )
    {
        /* lbfgs_atlas_transform_direction, Solver_Kernels.cpp:182-242.  The reference adds the
           per-spin terms into rho[n] from parallel threads without synchronisation; here the
           terms are summed deterministically and added once. */
        for( int n = 0; n < ORC_LBFGS_MEMORY; ++n )
        {
            struct atlas_ctx a = { spins, st->a3, st->da[n], st->dg[n] };
            double add         = det_sum( ns, term_atlas_rho, &( struct sum_ctx ){ NULL, NULL, NULL, (const double *)&a } );
            st->rho[n]         = 1 / st->rho[n];
            st->rho[n]         = st->rho[n] + add;
            st->rho[n]         = 1 / st->rho[n];
        }
        for( int i = 0; i < ns; ++i )
        {
            const double sz = spins[3 * i + 2];
            if( sz * st->a3[i] < 0 )
            {
                st->a3[i]           = sz > 0 ? 1 : -1;
                const double factor = ( 1 - st->a3[i] * sz ) / ( 1 + st->a3[i] * sz );
                st->sd[2 * i] *= factor;
                st->sd[2 * i + 1] *= factor;
                st->grad_pr[2 * i] *= factor;
                st->grad_pr[2 * i + 1] *= factor;
                for( int n = 0; n < ORC_LBFGS_MEMORY; ++n )
                {
                    st->da[n][2 * i] *= factor;
                    st->da[n][2 * i + 1] *= factor;
                    st->dg[n][2 * i] *= factor;
                    st->dg[n][2 * i + 1] *= factor;
                }
            }
        }
    }
}

/* One LLG iteration of a direct-minimisation solver; fills s->f1 (force) and s->fv1 (the
   virtual force Hook_Post_Iteration projects).  Returns the energy of the force evaluation. */
static double min_iteration( orc_sys * s, int solver, const orc_llg * par, double * spins, min_state * st )
{
    double E = calc_force( s, spins, s->f1 );
    calc_force_virtual( s, par, solver, spins, s->f1, s->fv1 );
    min_update( solver, par->dt, 1, s->nos, spins, s->f1, st, s->tmp );
    return E;
}

/* One LLG iteration of a dynamics solver (Method_LLG.cpp:66-301 + Solver_*.hpp) */
static double llg_step( orc_sys * s, int solver, const orc_llg * par, double * spins, uint64_t step, int it )
{
    const int nos = s->nos;
    double E      = 0;
    prepare_thermal( s, par, step, it );
    if( solver == ORC_SOLVER_DEPONDT )
    {
        /* Solver_Depondt.hpp:29-77 */
        E = calc_force( s, spins, s->f1 );
        calc_force_virtual( s, par, solver, spins, s->f1, s->fv1 );
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
        {
            double axis[3] = { 1 * s->fv1[3 * i], 1 * s->fv1[3 * i + 1], 1 * s->fv1[3 * i + 2] };
            double angle   = sqrt( v3_sqnorm( s->fv1 + 3 * i ) );
            v3_normalize( axis );
            orc_rotate( spins + 3 * i, axis, angle, s->sp + 3 * i );
        }
        E = calc_force( s, s->sp, s->f2 );
        calc_force_virtual( s, par, solver, s->sp, s->f2, s->fv2 );
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
        {
            double t[3];
            for( int q = 0; q < 3; ++q )
                t[q] = 0.5 * s->fv1[3 * i + q];
            for( int q = 0; q < 3; ++q )
                t[q] += 0.5 * s->fv2[3 * i + q];
            double angle = sqrt( v3_sqnorm( t ) );
            v3_normalize( t );
            orc_rotate( spins + 3 * i, t, angle, spins + 3 * i );
        }
    }
    else if( solver == ORC_SOLVER_HEUN )
    {
        /* Solver_Heun.hpp:30-83 */
        E = calc_force( s, spins, s->f1 );
        calc_force_virtual( s, par, solver, spins, s->f1, s->fv1 );
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
        {
            double * t  = s->tmp + 3 * i;
            double * pr = s->sp + 3 * i;
            neg_cross( spins + 3 * i, s->fv1 + 3 * i, t );
            for( int q = 0; q < 3; ++q )
                pr[q] = 1 * spins[3 * i + q];
            for( int q = 0; q < 3; ++q )
                pr[q] += 1 * t[q];
            v3_normalize( pr );
        }
        E = calc_force( s, s->sp, s->f2 );
        calc_force_virtual( s, par, solver, s->sp, s->f2, s->fv2 );
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
        {
            double * t = s->tmp + 3 * i;
            double t1[3];
            for( int q = 0; q < 3; ++q )
                t[q] *= 0.5;
            for( int q = 0; q < 3; ++q )
                t[q] += 1 * spins[3 * i + q];
            neg_cross( s->sp + 3 * i, s->fv2 + 3 * i, t1 );
            for( int q = 0; q < 3; ++q )
                t[q] += 0.5 * t1[q];
            v3_normalize( t );
            for( int q = 0; q < 3; ++q )
                spins[3 * i + q] = t[q];
        }
    }
    else if( solver == ORC_SOLVER_SIB )
    {
        /* Solver_SIB.hpp:22-52 */
        E = calc_force( s, spins, s->f1 );
        calc_force_virtual( s, par, solver, spins, s->f1, s->fv1 );
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
        {
            double * pr = s->sp + 3 * i;
            orc_sib_transform( spins + 3 * i, s->fv1 + 3 * i, pr );
            for( int q = 0; q < 3; ++q )
                pr[q] += 1 * spins[3 * i + q];
            for( int q = 0; q < 3; ++q )
                pr[q] *= 0.5;
        }
        E = calc_force( s, s->sp, s->f2 );
        calc_force_virtual( s, par, solver, s->sp, s->f2, s->fv2 );
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
            orc_sib_transform( spins + 3 * i, s->fv2 + 3 * i, spins + 3 * i );
    }
    else if( solver == ORC_SOLVER_RK4 )
    {
        /* Solver_RK4.hpp:41-147 */
        double * ks[4] = { s->k1, s->k2, s->k3, s->k4 };
        const double cpred[3] = { 0.5, 0.5, 1.0 };
        E = calc_force( s, spins, s->f1 );
        calc_force_virtual( s, par, solver, spins, s->f1, s->fv1 );
        const double * conf = spins;
        const double * fv   = s->fv1;
        for( int stage = 0; stage < 3; ++stage )
        {
            double * k = ks[stage];
#pragma omp parallel for
            for( int i = 0; i < nos; ++i )
            {
                neg_cross( conf + 3 * i, fv + 3 * i, k + 3 * i );
                double pr[3];
                for( int q = 0; q < 3; ++q )
                    pr[q] = 1 * spins[3 * i + q];
                for( int q = 0; q < 3; ++q )
                    pr[q] += cpred[stage] * k[3 * i + q];
                v3_normalize( pr );
                for( int q = 0; q < 3; ++q )
                    s->tmp[3 * i + q] = pr[q];
            }
            memcpy( s->sp, s->tmp, sizeof( double ) * 3 * (size_t)nos );
            E = calc_force( s, s->sp, s->f2 );
            calc_force_virtual( s, par, solver, s->sp, s->f2, s->fv2 );
            conf = s->sp;
            fv   = s->fv2;
        }
#pragma omp parallel for
        for( int i = 0; i < nos; ++i )
        {
            neg_cross( s->sp + 3 * i, s->fv2 + 3 * i, s->k4 + 3 * i );
            double t[3];
            for( int q = 0; q < 3; ++q )
                t[q] = 1 * spins[3 * i + q];
            for( int q = 0; q < 3; ++q )
                t[q] += 1.0 / 6.0 * s->k1[3 * i + q];
            for( int q = 0; q < 3; ++q )
                t[q] += 1.0 / 3.0 * s->k2[3 * i + q];
            for( int q = 0; q < 3; ++q )
                t[q] += 1.0 / 3.0 * s->k3[3 * i + q];
            for( int q = 0; q < 3; ++q )
                t[q] += 1.0 / 6.0 * s->k4[3 * i + q];
            v3_normalize( t );
            for( int q = 0; q < 3; ++q )
                spins[3 * i + q] = t[q];
        }
    }
    return E;
}

/* Hook_Post_Iteration, Method_LLG.cpp:247-301 and Method_Solver.hpp:224-230: max torque of
   the virtual force s->fv1 projected on the new spins */
static double llg_max_torque( const orc_sys * s, const double * spins )
{
    const int nos = s->nos;
    double m      = 0;
#pragma omp parallel for reduction( max : m )
    for( int i = 0; i < nos; ++i )
    {
        const double * f = s->fv1 + 3 * i;
        const double * v = spins + 3 * i;
        double d         = v3_dot( f, v );
        double pr[3];
        for( int q = 0; q < 3; ++q )
            pr[q] = f[q] - d * v[q];
        double n2 = v3_sqnorm( pr );
        if( n2 > m )
            m = n2;
    }
    return sqrt( m );
}

static long llg_loop( orc_sys * s, int solver, const orc_llg * par, double * spins, long n_steps, uint64_t step0,
                      double force_convergence, int check, double * energy, double * max_torque, double * eff_field )
{
    const int nos   = s->nos;
    double E        = 0;
    min_state * st  = is_minimiser( solver ) ? min_state_create( solver, nos, spins ) : NULL;
    long done       = 0;
    while( done < n_steps )
    {
        if( st )
            E = min_iteration( s, solver, par, spins, st );
        else
            E = llg_step( s, solver, par, spins, step0 + (uint64_t)done, (int)done );
        ++done;
        if( check && llg_max_torque( s, spins ) < force_convergence )
            break;
    }
    min_state_destroy( st );
    if( energy )
        *energy = E;
    if( max_torque && done > 0 )
        *max_torque = llg_max_torque( s, spins );
    if( eff_field && done > 0 )
    {
        for( int i = 0; i < nos; ++i )
        {
            const double * f = s->f1 + 3 * i;
            const double * v = spins + 3 * i;
            double d         = v3_dot( f, v );
            for( int q = 0; q < 3; ++q )
                eff_field[3 * i + q] = f[q] - d * v[q];
        }
    }
    return done;
}

void orc_llg_iterate( orc_sys * s, int solver, const orc_llg * par, double * spins, int n_steps, uint64_t step0,
                      double * energy, double * max_torque, double * eff_field )
{
    llg_loop( s, solver, par, spins, n_steps, step0, 0.0, 0, energy, max_torque, eff_field );
}

long orc_llg_run( orc_sys * s, int solver, const orc_llg * par, double * spins, long n_iterations,
                  double force_convergence, double * energy, double * max_torque, double * eff_field )
{
    return llg_loop( s, solver, par, spins, n_iterations, 0, force_convergence, 1, energy, max_torque, eff_field );
}

/* ======================================================================
 * GNEB (core/src/engine/Method_GNEB.cpp, Manifoldmath.cpp, Solver_*.hpp)
 * ====================================================================== */

/* Vectormath::dot over a field of n3/3 vectors (serial sum in index order) */
static double vf_dot( const double * a, const double * b, long n3 )
{
    double s = 0;
    for( long i = 0; i < n3; i += 3 )
        s += a[i] * b[i] + a[i + 1] * b[i + 1] + a[i + 2] * b[i + 2];
    return s;
}
/* Manifoldmath::normalize, Manifoldmath.cpp:220-229: vf *= 1/|vf| */
static void vf_normalize( double * a, long n3 )
{
    double sc = 1.0 / sqrt( vf_dot( a, a, n3 ) );
    for( long t = 0; t < n3; ++t )
        a[t] *= sc;
}
/* Manifoldmath::project_tangential, Manifoldmath.cpp:58-63 */
static void vf_project_tangential( double * v1, const double * v2, long n3 )
{
    for( long i = 0; i < n3; i += 3 )
    {
        double d = v1[i] * v2[i] + v1[i + 1] * v2[i + 1] + v1[i + 2] * v2[i + 2];
        for( int q = 0; q < 3; ++q )
            v1[i + q] -= d * v2[i + q];
    }
}
/* Vectormath::angle, Vectormath.cpp:434-441 */
static double v3_angle( const double * a, const double * b )
{
    double r = v3_dot( a, b );
    r        = fmax( -1.0, fmin( 1.0, r ) );
    return acos( r );
}
/* Manifoldmath::dist_geodesic, Manifoldmath.cpp:65-73 */
static double dist_geodesic( const double * a, const double * b, long n3 )
{
    double d = 0;
    for( long i = 0; i < n3; i += 3 )
    {
        double x = v3_angle( a + i, b + i );
        d += x * x;
    }
    return sqrt( d );
}

/* Geodesic_Tangent, Manifoldmath.cpp:78-108 */
static void geodesic_tangent( double * tangent, const double * minus, const double * plus, const double * mid, long n3 )
{
    const double ex[3] = { 1, 0, 0 }, ey[3] = { 0, 1, 0 };
    const double eps   = 1e-15;
    for( long i = 0; i < n3; i += 3 )
    {
        double axis[3];
        v3_cross( plus + i, minus + i, axis );
        if( fabs( v3_dot( minus + i, plus + i ) + 1 ) < eps )
        {
            const double * a = ( fabs( v3_dot( mid + i, ex ) - 1 ) > eps ) ? ex : ey;
            axis[0] = a[0], axis[1] = a[1], axis[2] = a[2];
        }
        v3_cross( mid + i, axis, tangent + i );
    }
    vf_normalize( tangent, n3 );
}

/* Tangents, Manifoldmath.cpp:113-209 */
static void gneb_tangents( double * const * conf, const double * E, double ** tangents, int noi, long n3, double * tp,
                           double * tm )
{
    for( int img = 0; img < noi; ++img )
    {
        const double * image = conf[img];
        double * tau         = tangents[img];
        if( img == 0 )
            geodesic_tangent( tau, image, conf[1], image, n3 );
        else if( img == noi - 1 )
            geodesic_tangent( tau, conf[img - 1], image, image, n3 );
        else
        {
            const double * plus  = conf[img + 1];
            const double * minus = conf[img - 1];
            double E_mid = E[img], E_plus = E[img + 1], E_minus = E[img - 1];
            for( long t = 0; t < n3; ++t )
            {
                tp[t] = 1 * plus[t];
                tp[t] += -1 * image[t];
                tm[t] = 1 * image[t];
                tm[t] += -1 * minus[t];
            }
            if( ( E_plus < E_mid && E_mid > E_minus ) || ( E_plus > E_mid && E_mid < E_minus ) )
            {
                double E_max = fmax( fabs( E_plus - E_mid ), fabs( E_minus - E_mid ) );
                double E_min = fmin( fabs( E_plus - E_mid ), fabs( E_minus - E_mid ) );
                double cp = E_plus > E_minus ? E_max : E_min;
                double cm = E_plus > E_minus ? E_min : E_max;
                for( long t = 0; t < n3; ++t )
                {
                    tau[t] = cp * tp[t];
                    tau[t] += cm * tm[t];
                }
            }
            else if( E_plus > E_mid && E_mid > E_minus )
            {
                for( long t = 0; t < n3; ++t )
                    tau[t] = 1 * tp[t];
            }
            else if( E_plus < E_mid && E_mid < E_minus )
            {
                for( long t = 0; t < n3; ++t )
                    tau[t] = tm[t];
            }
            else
            {
                for( long t = 0; t < n3; ++t )
                {
                    tau[t] = 1 * tp[t];
                    tau[t] += 1 * tm[t];
                }
            }
            vf_project_tangential( tau, image, n3 );
            vf_normalize( tau, n3 );
        }
    }
}

/* Cubic_Hermite_Spline::Interpolate, core/src/utility/Cubic_Hermite_Spline.cpp:10-47.
   x, p, m of length n; out_x, out_p of length n + (n-1)*n_interp. */
void orc_cubic_hermite( const double * x, const double * p, const double * m, int n, int n_interp, double * out_x,
                        double * out_p )
{
    for( int i = 0; i < n - 1; ++i )
    {
        double x0 = x[i], x1 = x[i + 1], p0 = p[i], p1 = p[i + 1], m0 = m[i], m1 = m[i + 1];
        for( int j = 0; j < n_interp + 1; ++j )
        {
            double t   = j / (double)( n_interp + 1 );
            double h00 = 2 * pow( t, 3 ) - 3 * pow( t, 2 ) + 1;
            double h10 = -2 * pow( t, 3 ) + 3 * pow( t, 2 );
            double h01 = pow( t, 3 ) - 2 * pow( t, 2 ) + t;
            double h11 = pow( t, 3 ) - pow( t, 2 );
            int idx    = i * ( n_interp + 1 ) + j;
            out_x[idx] = x0 + t * ( x1 - x0 );
            out_p[idx] = h00 * p0 + h10 * p1 + h01 * m0 * ( x0 - x1 ) + h11 * m1 * ( x0 - x1 );
        }
    }
    int last    = n + ( n - 1 ) * n_interp - 1;
    out_x[last] = x[n - 1];
    out_p[last] = p[n - 1];
}

/* Rotation matrix about a unit axis (Rodrigues), as Eigen::AngleAxis::toRotationMatrix */
static void rotation_matrix( const double * axis, double angle, double R[3][3] )
{
    double s = sin( angle ), c = cos( angle );
    double sa[3] = { s * axis[0], s * axis[1], s * axis[2] };
    double ca[3] = { ( 1 - c ) * axis[0], ( 1 - c ) * axis[1], ( 1 - c ) * axis[2] };
    double t;
    t       = ca[0] * axis[1];
    R[0][1] = t - sa[2];
    R[1][0] = t + sa[2];
    t       = ca[0] * axis[2];
    R[0][2] = t + sa[1];
    R[2][0] = t - sa[1];
    t       = ca[1] * axis[2];
    R[1][2] = t - sa[0];
    R[2][1] = t + sa[0];
    for( int q = 0; q < 3; ++q )
        R[q][q] = ca[q] * axis[q] + c;
}

typedef struct gneb_ws
{
    int noi;
    long n3;
    double **F_gradient, **F_total, **F_spring, **tangents, **eff;
    double *f_shrink, *F_tl, *F_tr, *tp, *tm, *lengths, *dE_dRx;
    double * Rx;
    double * E;
    int stop; /* geodesic distance zero */
} gneb_ws;

static double ** alloc_fields( int noi, long n3 )
{
    double ** f = (double **)malloc( sizeof( double * ) * noi );
    for( int i = 0; i < noi; ++i )
        f[i] = (double *)calloc( (size_t)n3, sizeof( double ) );
    return f;
}
static void free_fields( double ** f, int noi )
{
    if( !f )
        return;
    for( int i = 0; i < noi; ++i )
        free( f[i] );
    free( f );
}

/* Method_GNEB::Calculate_Force, Method_GNEB.cpp:88-357 */
static void gneb_calc_force( orc_sys * s, const orc_gneb * par, const int * image_type, gneb_ws * w,
                             double * const * conf, double * const * chain_spins, double ** forces )
{
    int noi  = w->noi;
    long n3  = w->n3;
    long nos = n3 / 3;
    for( int img = 0; img < noi; ++img )
    {
        orc_gradient_and_energy( s, conf[img], w->eff[img], &w->E[img] );
        for( long t = 0; t < n3; ++t )
        {
            w->eff[img][t] *= -1;
            w->F_gradient[img][t] = w->eff[img][t];
        }
        if( img > 0 )
        {
            w->Rx[img] = w->Rx[img - 1] + dist_geodesic( conf[img], conf[img - 1], n3 );
            if( w->Rx[img] - w->Rx[img - 1] < 1e-10 )
            {
                w->stop = 1;
                return;
            }
        }
    }
    gneb_tangents( conf, w->E, w->tangents, noi, n3, w->tp, w->tm );

    for( int i = 0; i < noi; ++i )
        w->lengths[i] = 0;
    if( par->spring_force_ratio > 0 )
    {
        double ratio_E  = fmin( 1.0, par->spring_force_ratio );
        double ratio_Rx = 1 - ratio_E;
        for( int i = 0; i < noi; ++i )
            w->dE_dRx[i] = vf_dot( w->eff[i], w->tangents[i], n3 );
        int n_interp = 20;
        int npts     = noi + ( noi - 1 ) * n_interp;
        double * ix  = (double *)malloc( sizeof( double ) * (size_t)( npts > 0 ? npts : 1 ) );
        double * ip  = (double *)malloc( sizeof( double ) * (size_t)( npts > 0 ? npts : 1 ) );
        orc_cubic_hermite( w->Rx, w->E, w->dE_dRx, noi, n_interp, ix, ip );
        int imax = 0, imin = 0, emax = 0, emin = 0;
        for( int k = 1; k < npts; ++k )
        {
            if( ix[k] > ix[imax] )
                imax = k;
            if( ix[k] < ix[imin] )
                imin = k;
            if( ip[k] > ip[emax] )
                emax = k;
            if( ip[k] < ip[emin] )
                emin = k;
        }
        double range_Rx = ix[imax] - ix[imin];
        double range_E  = ip[emax] - ip[emin];
        for( int idx_image = 1; idx_image < noi; ++idx_image )
        {
            for( int i = 1; i <= n_interp; ++i )
            {
                int idx    = ( idx_image - 1 ) * ( n_interp + 1 ) + i;
                double dRx = ratio_Rx * ( ix[idx] - ix[idx - 1] ) / range_Rx;
                double dE  = ratio_E * ( ip[idx] - ip[idx - 1] ) / range_E;
                w->lengths[idx_image] += sqrt( dRx * dRx + dE * dE );
            }
            w->lengths[idx_image] *= range_Rx;
        }
        free( ix );
        free( ip );
    }

    for( int img = 1; img < noi - 1; ++img )
    {
        const double * image = conf[img];
        double * Fg          = w->F_gradient[img];
        double * tau         = w->tangents[img];
        double * Ft          = w->F_total[img];
        vf_project_tangential( Fg, image, n3 );
        if( image_type[img] == 1 ) /* climbing: invert_parallel, Manifoldmath.cpp:41-47 */
        {
            double x = vf_dot( Fg, tau, n3 );
            for( long t = 0; t < n3; ++t )
                Fg[t] -= 2 * x * tau[t];
            memcpy( Ft, Fg, sizeof( double ) * n3 );
        }
        else if( image_type[img] == 2 ) /* falling */
            memcpy( Ft, Fg, sizeof( double ) * n3 );
        else if( image_type[img] == 0 ) /* normal */
        {
            double x = vf_dot( Fg, tau, n3 ); /* project_orthogonal, Manifoldmath.cpp:32-39 */
            for( long t = 0; t < n3; ++t )
                Fg[t] -= x * tau[t];
            if( par->path_shortening_constant > 0 )
            {
                double * tp = w->tp;
                double * tm = w->tm;
                for( long t = 0; t < n3; ++t )
                {
                    tp[t] = 1 * chain_spins[img + 1][t];
                    tp[t] += -1 * chain_spins[img][t];
                    tm[t] = 1 * chain_spins[img][t];
                    tm[t] += -1 * chain_spins[img - 1][t];
                }
                vf_normalize( tp, n3 );
                vf_normalize( tm, n3 );
                for( long t = 0; t < n3; ++t )
                {
                    w->f_shrink[t] = 1 * tp[t];
                    w->f_shrink[t] += -1 * tm[t];
                }
                for( long t = 0; t < n3; ++t )
                    tp[t] = 1 * Fg[t];
                double gradnorm = sqrt( vf_dot( tp, tp, n3 ) );
                for( long t = 0; t < n3; ++t )
                    tp[t] *= 1.0 / gradnorm;
                double x1 = vf_dot( w->f_shrink, tp, n3 );
                for( long t = 0; t < n3; ++t )
                    w->f_shrink[t] -= x1 * tp[t];
                double x2 = vf_dot( w->f_shrink, tau, n3 );
                for( long t = 0; t < n3; ++t )
                    w->f_shrink[t] -= x2 * tau[t];
                vf_normalize( w->f_shrink, n3 );
                double scalefactor = fmax( gradnorm, nos * par->path_shortening_constant );
                for( long t = 0; t < n3; ++t )
                    w->f_shrink[t] *= scalefactor;
            }
            double d;
            if( par->spring_force_ratio > 0 )
                d = par->spring_constant * ( w->lengths[img + 1] - w->lengths[img] );
            else
                d = par->spring_constant * ( w->Rx[img + 1] - 2 * w->Rx[img] + w->Rx[img - 1] );
            for( long t = 0; t < n3; ++t )
                w->F_spring[img][t] = d * tau[t];
            for( long t = 0; t < n3; ++t )
            {
                Ft[t] = 1 * Fg[t];
                Ft[t] += 1 * w->F_spring[img][t];
            }
            if( par->path_shortening_constant > 0 )
                for( long t = 0; t < n3; ++t )
                    Ft[t] += 1 * w->f_shrink[t];
        }
        else
            memset( Ft, 0, sizeof( double ) * n3 );
        for( long t = 0; t < n3; ++t )
            forces[img][t] = 1 * Ft[t];
    }

    if( par->moving_endpoints )
    {
        vf_project_tangential( w->F_gradient[0], conf[0], n3 );
        vf_project_tangential( w->F_gradient[noi - 1], conf[noi - 1], n3 );
        if( par->translating_endpoints )
        {
            const double * sl = chain_spins[0];
            const double * sr = chain_spins[noi - 1];
            const double * gl = w->F_gradient[0];
            const double * gr = w->F_gradient[noi - 1];
            for( long i = 0; i < n3; i += 3 )
            {
                double axis[3];
                v3_cross( sl + i, sr + i, axis );
                double angle = acos( v3_dot( sl + i, sr + i ) );
                v3_normalize( axis );
                double R[3][3];
                rotation_matrix( axis, angle, R );
                if( fabs( v3_dot( sl + i, sr + i ) ) >= 1.0 )
                    for( int a = 0; a < 3; ++a )
                        for( int b = 0; b < 3; ++b )
                            R[a][b] = a == b ? 1.0 : 0.0;
                for( int a = 0; a < 3; ++a )
                {
                    double rr = R[a][0] * gr[i] + R[a][1] * gr[i + 1] + R[a][2] * gr[i + 2];
                    double rl = R[0][a] * gl[i] + R[1][a] * gl[i + 1] + R[2][a] * gl[i + 2];
                    w->F_tl[i + a] = -0.5 * ( gl[i + a] + rr );
                    w->F_tr[i + a] = -0.5 * ( rl + gr[i + a] );
                }
            }
            double pl = vf_dot( w->F_tl, w->tangents[0], n3 );
            for( long t = 0; t < n3; ++t )
                w->F_tl[t] = pl * w->tangents[0][t];
            double pr = vf_dot( w->F_tr, w->tangents[noi - 1], n3 );
            for( long t = 0; t < n3; ++t )
                w->F_tr[t] = pr * w->tangents[noi - 1][t];
        }
        double rotational_coeff = 1.0;
        if( par->escape_first )
        {
            double proj_left  = vf_dot( w->F_gradient[0], w->tangents[0], n3 );
            double proj_right = vf_dot( w->F_gradient[noi - 1], w->tangents[noi - 1], n3 );
            if( proj_left > proj_right )
                rotational_coeff = 0.0;
        }
        for( int e = 0; e < 2; ++e )
        {
            int img           = e == 0 ? 0 : noi - 1;
            double delta_Rx0  = e == 0 ? par->delta_Rx_left : par->delta_Rx_right;
            double delta_Rx   = e == 0 ? w->Rx[1] - w->Rx[0] : w->Rx[noi - 1] - w->Rx[noi - 2];
            double k          = ( e == 0 ? 1.0 : -1.0 ) * par->spring_constant;
            double projection = vf_dot( w->F_gradient[img], w->tangents[img], n3 );
            double tc         = k * ( delta_Rx - delta_Rx0 );
            const double * Ftr = e == 0 ? w->F_tl : w->F_tr;
            for( long t = 0; t < n3; ++t )
            {
                forces[img][t] = rotational_coeff * ( w->F_gradient[img][t] - projection * w->tangents[img][t] )
                                 + tc * w->tangents[img][t] + Ftr[t];
                w->F_total[img][t] = forces[img][t];
            }
        }
    }
}

/* Method_GNEB::Calculate_Force_Virtual, Method_GNEB.cpp:360-391 */
static void gneb_force_virtual( const orc_gneb * par, int noi, long n3, double * const * conf, double ** forces,
                                double ** fv )
{
    double dtg = par->dt * ORC_GAMMA / ORC_MU_B;
    for( int i = 0; i < noi; ++i )
    {
        if( !par->moving_endpoints && ( i == 0 || i == noi - 1 ) )
            continue;
        for( long t = 0; t < n3; t += 3 )
        {
            double c[3];
            v3_cross( conf[i] + t, forces[i] + t, c );
            for( int q = 0; q < 3; ++q )
                fv[i][t + q] = dtg * c[q];
        }
    }
}

long orc_gneb_run( orc_sys * s, int solver, const orc_gneb * par, int noi, const int * image_type, double * spins,
                   long n_iterations, int check_convergence, double * energies, double * Rx, double * max_torque_all,
                   double * eff_field, double * Rx_interp, double * E_interp )
{
    if( noi < 2 )
        return 0;
    long n3 = 3L * s->nos;
    gneb_ws w;
    memset( &w, 0, sizeof( w ) );
    w.noi        = noi;
    w.n3         = n3;
    w.F_gradient = alloc_fields( noi, n3 );
    w.F_total    = alloc_fields( noi, n3 );
    w.F_spring   = alloc_fields( noi, n3 );
    w.tangents   = alloc_fields( noi, n3 );
    w.eff        = alloc_fields( noi, n3 );
    w.f_shrink   = (double *)calloc( (size_t)n3, sizeof( double ) );
    w.F_tl       = (double *)calloc( (size_t)n3, sizeof( double ) );
    w.F_tr       = (double *)calloc( (size_t)n3, sizeof( double ) );
    w.tp         = (double *)calloc( (size_t)n3, sizeof( double ) );
    w.tm         = (double *)calloc( (size_t)n3, sizeof( double ) );
    w.lengths    = (double *)calloc( (size_t)noi, sizeof( double ) );
    w.dE_dRx     = (double *)calloc( (size_t)noi, sizeof( double ) );
    w.Rx         = (double *)calloc( (size_t)noi, sizeof( double ) );
    w.E          = (double *)calloc( (size_t)noi, sizeof( double ) );

    double ** conf  = (double **)malloc( sizeof( double * ) * noi );
    for( int i = 0; i < noi; ++i )
        conf[i] = spins + (size_t)i * n3;
    double ** forces = alloc_fields( noi, n3 );
    double ** fv     = alloc_fields( noi, n3 );
    double ** pred   = alloc_fields( noi, n3 );
    double ** fpred  = alloc_fields( noi, n3 );
    double ** fvpred = alloc_fields( noi, n3 );
    double ** tmp    = alloc_fields( noi, n3 );
    double ** vel    = alloc_fields( noi, n3 );
    double ** fprev  = alloc_fields( noi, n3 );
    double ** k1 = alloc_fields( noi, n3 ), **k2 = alloc_fields( noi, n3 ), **k3 = alloc_fields( noi, n3 );
    double * mt = (double *)calloc( (size_t)noi, sizeof( double ) );
    double * proj = (double *)malloc( sizeof( double ) * (size_t)noi );
    double * fn2  = (double *)malloc( sizeof( double ) * (size_t)noi );
    const double m = 1.0;
    /* VP_OSO, LBFGS_OSO, LBFGS_Atlas: solver state over the whole chain */
    const int oso_atlas = solver == ORC_SOLVER_VP_OSO || solver == ORC_SOLVER_LBFGS_OSO
                          || solver == ORC_SOLVER_LBFGS_ATLAS;
    min_state * mst = oso_atlas ? min_state_create( solver, noi * s->nos, spins ) : NULL;
    double * Fc     = oso_atlas ? (double *)calloc( (size_t)noi * n3, sizeof( double ) ) : NULL;
    double * gtmp   = oso_atlas ? (double *)calloc( (size_t)noi * n3, sizeof( double ) ) : NULL;

    double max_torque = par->force_convergence + 1.0;
    long iteration    = 0;
    for( iteration = 0; iteration < n_iterations; ++iteration )
    {
        if( check_convergence && max_torque < par->force_convergence )
            break;
        if( solver == ORC_SOLVER_VP )
        {
            /* Solver_VP.hpp:29-114 */
            for( int i = 0; i < noi; ++i )
                memcpy( fprev[i], forces[i], sizeof( double ) * n3 );
            gneb_calc_force( s, par, image_type, &w, conf, conf, forces );
            if( w.stop )
                break;
            gneb_force_virtual( par, noi, n3, conf, forces, fv );
            double proj_full = 0, fn2_full = 0;
            for( int i = 0; i < noi; ++i )
            {
                for( long t = 0; t < n3; ++t )
                    vel[i][t] += 0.5 / m * ( fprev[i][t] + forces[i][t] );
                proj[i] = vf_dot( vel[i], forces[i], n3 );
                fn2[i]  = vf_dot( forces[i], forces[i], n3 );
            }
            for( int i = 0; i < noi; ++i )
            {
                proj_full += proj[i];
                fn2_full += fn2[i];
            }
            for( int i = 0; i < noi; ++i )
            {
                double dt    = par->dt;
                double ratio = proj_full / fn2_full;
                if( proj_full <= 0 )
                    memset( vel[i], 0, sizeof( double ) * n3 );
                else
                    for( long t = 0; t < n3; ++t )
                        vel[i][t] = forces[i][t] * ratio;
                for( long t = 0; t < n3; t += 3 )
                {
                    double c[3];
                    for( int q = 0; q < 3; ++q )
                        c[q] = conf[i][t + q] + dt * vel[i][t + q] + 0.5 / m * dt * forces[i][t + q];
                    v3_normalize( c );
                    for( int q = 0; q < 3; ++q )
                        conf[i][t + q] = c[q];
                }
            }
        }
        else if( oso_atlas )
        {
            /* Solver_VP_OSO.hpp:33-110, Solver_LBFGS_OSO.hpp:37-72, Solver_LBFGS_Atlas.hpp:44-101
               on the chain's forces F_total (Method_GNEB.cpp:88-357) */
            gneb_calc_force( s, par, image_type, &w, conf, conf, forces );
            if( w.stop )
                break;
            for( int i = 0; i < noi; ++i )
                memcpy( Fc + (size_t)i * n3, forces[i], sizeof( double ) * n3 );
            min_update( solver, par->dt, noi, s->nos, spins, Fc, mst, gtmp );
        }
        else if( solver == ORC_SOLVER_HEUN || solver == ORC_SOLVER_DEPONDT || solver == ORC_SOLVER_SIB )
        {
            gneb_calc_force( s, par, image_type, &w, conf, conf, forces );
            if( w.stop )
                break;
            gneb_force_virtual( par, noi, n3, conf, forces, fv );
            for( int i = 0; i < noi; ++i )
                for( long t = 0; t < n3; t += 3 )
                {
                    double * sp = pred[i] + t;
                    const double * x = conf[i] + t;
                    const double * f = fv[i] + t;
                    if( solver == ORC_SOLVER_HEUN )
                    {
                        neg_cross( x, f, tmp[i] + t );
                        for( int q = 0; q < 3; ++q )
                            sp[q] = 1 * x[q];
                        for( int q = 0; q < 3; ++q )
                            sp[q] += 1 * tmp[i][t + q];
                        v3_normalize( sp );
                    }
                    else if( solver == ORC_SOLVER_DEPONDT )
                    {
                        double axis[3] = { 1 * f[0], 1 * f[1], 1 * f[2] };
                        double angle   = sqrt( v3_sqnorm( f ) );
                        v3_normalize( axis );
                        orc_rotate( x, axis, angle, sp );
                    }
                    else
                    {
                        orc_sib_transform( x, f, sp );
                        for( int q = 0; q < 3; ++q )
                            sp[q] += 1 * x[q];
                        for( int q = 0; q < 3; ++q )
                            sp[q] *= 0.5;
                    }
                }
            gneb_calc_force( s, par, image_type, &w, pred, conf, fpred );
            if( w.stop )
                break;
            gneb_force_virtual( par, noi, n3, pred, fpred, fvpred );
            for( int i = 0; i < noi; ++i )
                for( long t = 0; t < n3; t += 3 )
                {
                    double * x = conf[i] + t;
                    if( solver == ORC_SOLVER_HEUN )
                    {
                        double * c = tmp[i] + t;
                        double t1[3];
                        for( int q = 0; q < 3; ++q )
                            c[q] *= 0.5;
                        for( int q = 0; q < 3; ++q )
                            c[q] += 1 * x[q];
                        neg_cross( pred[i] + t, fvpred[i] + t, t1 );
                        for( int q = 0; q < 3; ++q )
                            c[q] += 0.5 * t1[q];
                        v3_normalize( c );
                        for( int q = 0; q < 3; ++q )
                            x[q] = c[q];
                    }
                    else if( solver == ORC_SOLVER_DEPONDT )
                    {
                        double h[3];
                        for( int q = 0; q < 3; ++q )
                            h[q] = 0.5 * fv[i][t + q];
                        for( int q = 0; q < 3; ++q )
                            h[q] += 0.5 * fvpred[i][t + q];
                        double angle = sqrt( v3_sqnorm( h ) );
                        v3_normalize( h );
                        orc_rotate( x, h, angle, x );
                    }
                    else
                        orc_sib_transform( x, fvpred[i] + t, x );
                }
        }
        else if( solver == ORC_SOLVER_RK4 )
        {
            /* Solver_RK4.hpp:41-147 over all images */
            double ** ks[3] = { k1, k2, k3 };
            const double cpred[3] = { 0.5, 0.5, 1.0 };
            gneb_calc_force( s, par, image_type, &w, conf, conf, forces );
            if( w.stop )
                break;
            gneb_force_virtual( par, noi, n3, conf, forces, fv );
            double ** c_in  = conf;
            double ** fv_in = fv;
            for( int stage = 0; stage < 3 && !w.stop; ++stage )
            {
                for( int i = 0; i < noi; ++i )
                    for( long t = 0; t < n3; t += 3 )
                    {
                        neg_cross( c_in[i] + t, fv_in[i] + t, ks[stage][i] + t );
                        double p[3];
                        for( int q = 0; q < 3; ++q )
                            p[q] = 1 * conf[i][t + q];
                        for( int q = 0; q < 3; ++q )
                            p[q] += cpred[stage] * ks[stage][i][t + q];
                        v3_normalize( p );
                        for( int q = 0; q < 3; ++q )
                            tmp[i][t + q] = p[q];
                    }
                for( int i = 0; i < noi; ++i )
                    memcpy( pred[i], tmp[i], sizeof( double ) * n3 );
                gneb_calc_force( s, par, image_type, &w, pred, conf, fpred );
                gneb_force_virtual( par, noi, n3, pred, fpred, fvpred );
                c_in  = pred;
                fv_in = fvpred;
            }
            if( w.stop )
                break;
            for( int i = 0; i < noi; ++i )
                for( long t = 0; t < n3; t += 3 )
                {
                    double k4[3], c[3];
                    neg_cross( pred[i] + t, fvpred[i] + t, k4 );
                    for( int q = 0; q < 3; ++q )
                        c[q] = 1 * conf[i][t + q];
                    for( int q = 0; q < 3; ++q )
                        c[q] += 1.0 / 6.0 * k1[i][t + q];
                    for( int q = 0; q < 3; ++q )
                        c[q] += 1.0 / 3.0 * k2[i][t + q];
                    for( int q = 0; q < 3; ++q )
                        c[q] += 1.0 / 3.0 * k3[i][t + q];
                    for( int q = 0; q < 3; ++q )
                        c[q] += 1.0 / 6.0 * k4[q];
                    v3_normalize( c );
                    for( int q = 0; q < 3; ++q )
                        conf[i][t + q] = c[q];
                }
        }
        /* Hook_Post_Iteration, Method_GNEB.cpp:411-456 */
        max_torque = 0;
        for( int img = 0; img < noi; ++img )
        {
            vf_project_tangential( w.F_total[img], conf[img], n3 );
            double mx = 0;
            for( long t = 0; t < n3; t += 3 )
                mx = fmax( mx, v3_sqnorm( w.F_total[img] + t ) );
            mt[img] = sqrt( mx );
            if( mt[img] > max_torque )
                max_torque = mt[img];
            vf_project_tangential( forces[img], conf[img], n3 );
        }
    }

    if( energies )
        memcpy( energies, w.E, sizeof( double ) * noi );
    if( Rx )
        memcpy( Rx, w.Rx, sizeof( double ) * noi );
    if( max_torque_all )
        memcpy( max_torque_all, mt, sizeof( double ) * noi );
    if( eff_field )
        for( int i = 0; i < noi; ++i )
            memcpy( eff_field + (size_t)i * n3, w.eff[i], sizeof( double ) * n3 );
    if( Rx_interp && E_interp )
    {
        for( int i = 0; i < noi; ++i )
            w.dE_dRx[i] = vf_dot( w.eff[i], w.tangents[i], n3 );
        orc_cubic_hermite( w.Rx, w.E, w.dE_dRx, noi, par->n_E_interpolations, Rx_interp, E_interp );
    }
    long done = w.stop ? -( iteration + 1 ) : iteration;

    free_fields( w.F_gradient, noi );
    free_fields( w.F_total, noi );
    free_fields( w.F_spring, noi );
    free_fields( w.tangents, noi );
    free_fields( w.eff, noi );
    free( w.f_shrink );
    free( w.F_tl );
    free( w.F_tr );
    free( w.tp );
    free( w.tm );
    free( w.lengths );
    free( w.dE_dRx );
    free( w.Rx );
    free( w.E );
    free( conf );
    min_state_destroy( mst );
    free( Fc );
    free( gtmp );
    free_fields( forces, noi );
    free_fields( fv, noi );
    free_fields( pred, noi );
    free_fields( fpred, noi );
    free_fields( fvpred, noi );
    free_fields( tmp, noi );
    free_fields( vel, noi );
    free_fields( fprev, noi );
    free_fields( k1, noi );
    free_fields( k2, noi );
    free_fields( k3, noi );
    free( mt );
    free( proj );
    free( fn2 );
    return done;
}

/* Configuration_Chain::Homogeneous_Rotation, core/src/utility/Configuration_Chain.cpp:28-75.
   chain: noi images of nos spins (AoS); images idx_1 < img < idx_2 are overwritten. */
void orc_homogeneous_rotation( double * chain, int nos, int idx_1, int idx_2 )
{
    const double ex[3] = { 1, 0, 0 }, ey[3] = { 0, 1, 0 };
    long n3 = 3L * nos;
    for( int i = 0; i < nos; ++i )
    {
        const double * s1 = chain + idx_1 * n3 + 3 * i;
        const double * s2 = chain + idx_2 * n3 + 3 * i;
        double rot_angle  = v3_angle( s1, s2 );
        double axis[3];
        v3_cross( s1, s2, axis );
        v3_normalize( axis );
        if( fabs( rot_angle - ORC_PI ) < 1e-4 )
        {
            const double * a = ( fabs( v3_dot( s1, ex ) ) - 1 > 1e-4 ) ? ex : ey;
            axis[0] = a[0], axis[1] = a[1], axis[2] = a[2];
        }
        for( int img = idx_1 + 1; img < idx_2; ++img )
        {
            double * o = chain + img * n3 + 3 * i;
            if( rot_angle > 1e-8 )
            {
                double angle = rot_angle * (double)( img - idx_1 ) / (double)( idx_2 - idx_1 );
                orc_rotate( s1, axis, angle, o );
            }
            else
                o[0] = s1[0], o[1] = s1[1], o[2] = s1[2];
        }
    }
}
