// Device-resident LLG runner and DDI plan.
#include "gpu_engine.hpp"

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <atomic>
#include <thread>

namespace spirit_hip
{

void hip_check( hipError_t e, const char * what )
{
    if( e != hipSuccess )
        throw std::runtime_error( std::string( "HIP error " ) + hipGetErrorString( e ) + " in " + what );
}

DeviceBuffer::~DeviceBuffer()
{
    free();
}

void DeviceBuffer::free()
{
    if( p )
        (void)hipFree( p );
    p     = nullptr;
    bytes = 0;
}

void DeviceBuffer::resize( size_t n )
{
    if( n == bytes && p )
        return;
    free();
    if( n == 0 )
        return;
    HIPCHECK( hipMalloc( &p, n ) );
    bytes = n;
}

// ---------------------------------------------------------------------------
// FFT lengths
// ---------------------------------------------------------------------------
void FftLenHost::build( int n )
{
    if( n < 1 || n > 4096 )
        throw std::runtime_error( "DDI FFT: padded length " + std::to_string( n ) + " outside [1, 4096]" );
    dev.n        = n;
    dev.n_stages = 0;
    int m        = n;
    auto push    = [&]( int r ) {
        if( dev.n_stages >= MAX_FFT_STAGES )
            throw std::runtime_error( "DDI FFT: too many stages" );
        dev.radix[dev.n_stages++] = r;
        m /= r;
    };
    while( m % 8 == 0 )
        push( 8 );
    while( m % 4 == 0 )
        push( 4 );
    while( m % 2 == 0 )
        push( 2 );
    for( int r : { 3, 5, 7, 11, 13 } )
        while( m % r == 0 )
            push( r );
    if( m != 1 )
        throw std::runtime_error( "DDI FFT: padded length " + std::to_string( n )
                                  + " has a prime factor above 13 (supported radices 2,3,4,5,7,8,11,13)" );
    std::vector<double2> w( n );
    const long double twopi = 6.283185307179586476925286766559005768L;
    for( int e = 0; e < n; ++e )
    {
        long double a = -twopi * (long double)e / (long double)n;
        w[e]          = make_double2( (double)cosl( a ), (double)sinl( a ) );
    }
    W.resize( sizeof( double2 ) * n );
    HIPCHECK( hipMemcpy( W.as<void>(), w.data(), sizeof( double2 ) * n, hipMemcpyHostToDevice ) );
    dev.W   = W.as<double2>();
    dev.tw2 = nullptr;

    // Power-of-two plan: radix-8 stages then one radix-2/4 stage.  Stage s >= 1 has
    // Ns = 8^s and stores, at offset 8^s - 8, w_n^(r*jm*n/(Ns*R)) at jm*(R-1) + r-1.
    if( n >= 8 && ( n & ( n - 1 ) ) == 0 )
    {
        int lg = 0;
        while( ( 1 << lg ) < n )
            ++lg;
        const int n8  = lg / 3;
        const int rem = lg % 3;
        const int nst = n8 + ( rem ? 1 : 0 );
        std::vector<double2> t;
        int Ns = 8;
        for( int st = 1; st < nst; ++st, Ns *= 8 )
        {
            const int R = st < n8 ? 8 : ( rem == 1 ? 2 : 4 );
            if( (int)t.size() != Ns - 8 )
                throw std::runtime_error( "DDI FFT: twiddle table layout" );
            for( int jm = 0; jm < Ns; ++jm )
                for( int r = 1; r < R; ++r )
                    t.push_back( w[( (long long)r * jm * ( n / ( Ns * R ) ) ) % n] );
        }
        if( t.empty() )
            t.push_back( make_double2( 1.0, 0.0 ) );
        TW2.resize( sizeof( double2 ) * t.size() );
        HIPCHECK( hipMemcpy( TW2.as<void>(), t.data(), sizeof( double2 ) * t.size(), hipMemcpyHostToDevice ) );
        dev.tw2 = TW2.as<double2>();
    }
}

std::array<int, 3> ddi_padded_dims( const Geometry & g, const Hamiltonian & h )
{
    std::array<int, 3> p;
    for( int i = 0; i < 3; ++i )
    {
        p[i]     = g.n_cells[i];
        bool pad = g.n_cells[i] > 1 && ( h.ddi_bc[i] == 0 || h.ddi_pb_zero_padding );
        if( pad )
            p[i] *= 2;
    }
    return p;
}

size_t DdiPlan::bytes() const
{
    return X.size() + Dh.size() + b_pos.size() + b_pairs.size();
}

const char * DdiPlan::spectrum_format() const
{
    if( d.dh_real && d.sym_b && d.sym_c )
        return "real, mirror-symmetric in b and c";
    if( d.dh_real && d.sym_b )
        return "real, mirror-symmetric in b";
    if( d.dh_real && d.sym_c )
        return "real, mirror-symmetric in c";
    if( d.dh_real )
        return "real";
    return "complex";
}

// Cutoff pairs (from Hamiltonian::update_interactions) or the positions of the direct sum.
void DdiPlan::build_real_space( const Geometry & g, const Hamiltonian & h )
{
    if( g.nb > MAX_BASIS )
        throw std::runtime_error( "DDI: too many basis atoms" );
    rs        = DdiRsDev{};
    rs.method = h.ddi_radius < 0 ? DDI_RS_DIRECT : DDI_RS_CUTOFF;
    rs.nos    = g.nos;
    rs.nb     = g.nb;
    rs.Na     = g.n_cells[0];
    rs.Nb     = g.n_cells[1];
    rs.Nc     = g.n_cells[2];
    for( int k = 0; k < 3; ++k )
    {
        rs.bc[k]  = h.bc[k];
        rs.img[k] = h.bc[k] ? h.ddi_images[k] : 0;
        for( int c = 0; c < 3; ++c )
            rs.bv[k][c] = g.bravais[k][c];
    }
    rs.mult = mu_0 * mu_B * mu_B / ( 4 * Pi * 1e-30 );
    rs.lc   = g.lattice_constant;
    for( int b = 0; b < g.nb; ++b )
        rs.mu[b] = g.mu_basis[b];
    rs.div_nb = make_fast_div( g.nb );
    rs.div_Na = make_fast_div( g.n_cells[0] );
    rs.div_Nb = make_fast_div( g.n_cells[1] );
    if( rs.method == DDI_RS_DIRECT )
    {
        std::vector<double> soa( 3 * (size_t)g.nos );
        aos_to_soa( g.positions.data(), soa.data(), g.nos );
        b_pos.resize( sizeof( double ) * soa.size() );
        HIPCHECK( hipMemcpy( b_pos.as<void>(), soa.data(), sizeof( double ) * soa.size(), hipMemcpyHostToDevice ) );
        rs.pos = b_pos.as<double>();
    }
    else
    {
        std::vector<DdiPairDev> pv;
        for( size_t k = 0; k < h.ddi_pairs.size(); ++k )
        {
            // Gradient_DDI_Cutoff skips pairs of zero length (Hamiltonian_Heisenberg.cpp:887)
            if( !( h.ddi_dist[k] > 0.0 ) )
                continue;
            const Pair & p = h.ddi_pairs[k];
            DdiPairDev e{};
            e.i  = p.i;
            e.j  = p.j;
            e.da = p.t[0];
            e.db = p.t[1];
            e.dc = p.t[2];
            e.c  = rs.mult / std::pow( h.ddi_dist[k], 3.0 );
            for( int c = 0; c < 3; ++c )
                e.n[c] = h.ddi_n[k][c];
            pv.push_back( e );
        }
        rs.n_pairs = (int)pv.size();
        b_pairs.resize( sizeof( DdiPairDev ) * std::max<size_t>( 1, pv.size() ) );
        if( !pv.empty() )
            HIPCHECK( hipMemcpy( b_pairs.as<void>(), pv.data(), sizeof( DdiPairDev ) * pv.size(),
                                 hipMemcpyHostToDevice ) );
        rs.pairs = b_pairs.as<DdiPairDev>();
    }
}

void DdiPlan::build( const Geometry & g, const Hamiltonian & h, hipStream_t st )
{
    if( h.ddi_method == DDI_CUTOFF )
    {
        build_real_space( g, h );
        return;
    }
    rs     = DdiRsDev{};
    auto P = ddi_padded_dims( g, h );
    if( g.nb > MAX_BASIS )
        throw std::runtime_error( "DDI: too many basis atoms" );
    La.build( P[0] );
    Lb.build( P[1] );
    Lc.build( P[2] );
    d     = DdiDev{};
    d.Na  = g.n_cells[0];
    d.Nb  = g.n_cells[1];
    d.Nc  = g.n_cells[2];
    d.nb  = g.nb;
    d.Pa  = P[0];
    d.Pb  = P[1];
    d.Pc  = P[2];
    d.Ha  = P[0] / 2 + 1;
    d.Hs  = ( d.Ha + 7 ) / 8 * 8;
    d.M   = (double)P[0] * (double)P[1] * (double)P[2];
    {
        int dev_id = 0, lds = 0;
        HIPCHECK( hipGetDevice( &dev_id ) );
        HIPCHECK( hipDeviceGetAttribute( &lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id ) );
        d.lds_max = lds;
    }
    // Pass B reads what pass A wrote and pass A^-1 what pass B^-1 wrote, in the same
    // order: run B and A^-1 backwards so that they start on the rows written last, which
    // the 256 MB memory-side cache still holds (3.20 -> 3.15 ms per Depondt step at
    // 256x256x128).  SPIRIT_HIP_DDI_REVERSE = sum of DdiReverse bits overrides it (tuning).
    d.rev = DDI_REV_B_FWD | DDI_REV_A_INV;
    if( const char * r = std::getenv( "SPIRIT_HIP_DDI_REVERSE" ) )
        d.rev = std::atoi( r ) & 15;
    d.La  = La.dev;
    d.Lb  = Lb.dev;
    d.Lc  = Lc.dev;
    // inter_sublattice_lookup, Hamiltonian_Heisenberg.cpp:1415-1428 and 1545-1554
    int b_inter = -1;
    for( int b1 = 0; b1 < g.nb; ++b1 )
        for( int b2 = 0; b2 < g.nb; ++b2 )
        {
            if( b1 == b2 && b1 != 0 )
            {
                d.inter[b1 + b2 * g.nb] = 0;
                continue;
            }
            d.inter[b1 + b2 * g.nb] = ++b_inter;
        }
    d.n_inter = b_inter + 1;
    for( int b = 0; b < g.nb; ++b )
        d.mu[b] = g.mu_basis[b];

    if( const char * why = ddi_plan_problem( d ) )
        throw std::runtime_error( std::string( "DDI: " ) + why );

    const size_t nX = (size_t)3 * g.nb * d.Nc * (size_t)d.Pb * d.Hs;
    X.resize( sizeof( double2 ) * nX );
    d.X = X.as<double2>();
    // component signs of D under b -> -b and c -> -c for an orthogonal lattice
    // (xx, xy, xz, yy, yz, zz); used only where the spectrum shows them.
    const int sb[6] = { +1, -1, +1, +1, -1, +1 };
    const int sc[6] = { +1, +1, -1, +1, -1, +1 };
    for( int q = 0; q < 6; ++q )
    {
        d.sig_b[q] = sb[q];
        d.sig_c[q] = sc[q];
    }

    DipoleArgs a{};
    a.Na = d.Na, a.Nb = d.Nb, a.Nc = d.Nc, a.nb = d.nb;
    a.Pa = d.Pa, a.Pb = d.Pb, a.Pc = d.Pc;
    for( int i = 0; i < 3; ++i )
        a.img[i] = h.ddi_bc[i] == 0 ? 0 : h.ddi_images[i];
    a.lattice_constant = g.lattice_constant;
    for( int k = 0; k < 3; ++k )
        for( int c = 0; c < 3; ++c )
            a.bravais[k][c] = g.bravais[k][c];
    for( int b = 0; b < g.nb; ++b )
        for( int c = 0; c < 3; ++c )
            a.cell_atoms[b][c] = g.cell_atoms[b][c];

    const size_t M     = (size_t)d.Pa * d.Pb * d.Pc;
    const size_t nfull = (size_t)d.Pb * d.Ha * (size_t)d.Pc * 6 * d.n_inter;
    DeviceBuffer full( sizeof( double2 ) * nfull );
    {
        DeviceBuffer work( sizeof( double ) * 6 * d.n_inter * M );
        DeviceBuffer work2( sizeof( double2 ) * 6 * d.n_inter * (size_t)d.Pc * d.Pb * d.Hs );
        ddi_launch_kernel_spectrum( d, a, work.as<double>(), work2.as<double2>(), full.as<double2>(), st );
        HIPCHECK( hipGetLastError() );
        HIPCHECK( hipStreamSynchronize( st ) );
    }
    // Compressed storage where the spectrum allows it (decided from the computed values:
    // deviations below 1e-12 of the largest entry are rounding).
    DeviceBuffer sym_dev( sizeof( DhatSymmetry ) );
    ddi_launch_symmetry_check( d, full.as<double2>(), sym_dev.as<DhatSymmetry>(), st );
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipMemcpyAsync( &sym, sym_dev.as<void>(), sizeof( DhatSymmetry ), hipMemcpyDeviceToHost, st ) );
    HIPCHECK( hipStreamSynchronize( st ) );
    const double tol = 1e-12 * sym.max_abs;
    d.dh_real        = sym.max_imag <= tol;
    d.sym_b          = d.dh_real && sym.max_asym_b <= tol && d.Pb > 2;
    d.sym_c          = d.dh_real && sym.max_asym_c <= tol && d.Pc > 2;
    d.KB             = d.sym_b ? d.Pb / 2 + 1 : d.Pb;
    d.KC             = d.sym_c ? d.Pc / 2 + 1 : d.Pc;
    const size_t nDh = (size_t)d.KB * d.Ha * (size_t)d.KC * 6 * d.n_inter;
    Dh.resize( ( d.dh_real ? sizeof( double ) : sizeof( double2 ) ) * nDh );
    d.Dh   = d.dh_real ? nullptr : Dh.as<double2>();
    d.Dh_r = d.dh_real ? Dh.as<double>() : nullptr;
    ddi_launch_compress( d, full.as<double2>(), Dh.as<void>(), st );
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipStreamSynchronize( st ) );
    full.free();
    choose_pass_c( st );
    char msg[320];
    std::snprintf( msg, sizeof( msg ),
                   "DDI plan %dx%dx%d padded, kernel spectrum %s (|Im|/max %.2e, b-mirror %.2e, c-mirror %.2e), "
                   "pass C kernel %d/%d (%.3f ms)",
                   d.Pa, d.Pb, d.Pc, spectrum_format(), sym.max_imag / sym.max_abs, sym.max_asym_b / sym.max_abs,
                   sym.max_asym_c / sym.max_abs, d.c_kind, d.c_tk, c_ms );
    log_message( 5, msg );
}

// Pass C choices made in this process, by plan shape: the candidates do not round alike,
// so two plans of one shape (two images, a rebuilt plan) take the same kernel and give
// bit-identical fields, whatever the timing noise of a later measurement.
static std::mutex pass_c_mutex;
struct PassCChoice
{
    int kind, tk;
    double ms;
};
static std::map<std::vector<int>, PassCChoice> pass_c_chosen;

// Times every pass C kernel that can run the plan on a zero configuration (the run time
// does not depend on the values) and keeps the fastest.  The environment variable
// SPIRIT_HIP_DDI_PASS_C = generic | lds | reg | reg4 restricts the choice to that kernel where it
// can run the plan (for tests of each kernel and for tuning).
void DdiPlan::choose_pass_c( hipStream_t st )
{
    auto cand = ddi_pass_c_candidates( d );
    if( const char * force = std::getenv( "SPIRIT_HIP_DDI_PASS_C" ) )
    {
        const std::string f( force );
        const int want = f == "generic" ? DDI_C_GENERIC
                         : f == "lds"   ? DDI_C_P2_LDS
                         : f == "reg"   ? DDI_C_P2_REG
                         : f == "reg4"  ? DDI_C_P4_REG
                                        : -1;
        std::vector<std::pair<int, int>> keep;
        for( auto & c : cand )
            if( c.first == want )
                keep.push_back( c );
        if( keep.empty() && want == DDI_C_GENERIC )
            keep.push_back( { DDI_C_GENERIC, 0 } );
        if( keep.empty() )
            log_message( 2, "SPIRIT_HIP_DDI_PASS_C=" + f + " cannot run this DDI plan; timing every candidate" );
        else
            cand = keep;
    }
    d.c_kind  = cand[0].first;
    d.c_tk    = cand[0].second;
    c_ms      = 0;
    if( cand.size() == 1 )
        return;
    std::vector<int> key = { d.Pa, d.Pb, d.Pc, d.Hs, d.Na, d.Nb, d.Nc, d.nb, d.n_inter, d.dh_real, d.sym_b,
                             d.sym_c, d.KB, d.KC, d.lds_max };
    for( auto & c : cand )
        key.push_back( c.first ), key.push_back( c.second );
    {
        std::lock_guard<std::mutex> lk( pass_c_mutex );
        auto it = pass_c_chosen.find( key );
        if( it != pass_c_chosen.end() )
        {
            d.c_kind = it->second.kind;
            d.c_tk   = it->second.tk;
            c_ms     = it->second.ms;
            return;
        }
    }
    const size_t nos = (size_t)d.Na * d.Nb * d.Nc * d.nb;
    DeviceBuffer s0( sizeof( double ) * 3 * nos ), h0( sizeof( double ) * 3 * nos );
    HIPCHECK( hipMemsetAsync( s0.as<void>(), 0, s0.size(), st ) );
    hipEvent_t e0, e1;
    HIPCHECK( hipEventCreate( &e0 ) );
    HIPCHECK( hipEventCreate( &e1 ) );
    double best = 1e30;
    std::pair<int, int> pick = cand[0];
    for( auto & c : cand )
    {
        d.c_kind = c.first;
        d.c_tk   = c.second;
        ddi_launch_field( d, s0.as<double>(), h0.as<double>(), (int)nos, st );
        double t = 1e30;
        for( int rep = 0; rep < 3; ++rep )
        {
            ddi_launch_field( d, s0.as<double>(), h0.as<double>(), (int)nos, st, e0, e1 );
            HIPCHECK( hipGetLastError() );
            HIPCHECK( hipEventSynchronize( e1 ) );
            float ms = 0;
            HIPCHECK( hipEventElapsedTime( &ms, e0, e1 ) );
            t = std::min( t, (double)ms );
        }
        char msg[96];
        std::snprintf( msg, sizeof( msg ), "DDI pass C candidate %d/%d: %.4f ms", c.first, c.second, t );
        log_message( 6, msg );
        if( t < best )
        {
            best = t;
            pick = c;
        }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    d.c_kind = pick.first;
    d.c_tk   = pick.second;
    c_ms     = best;
    std::lock_guard<std::mutex> lk( pass_c_mutex );
    pass_c_chosen.emplace( key, PassCChoice{ pick.first, pick.second, best } );
}

void DdiPlan::field( const double * spins_soa, double * h_soa, int nos, hipStream_t st, hipEvent_t evc0,
                     hipEvent_t evc1 )
{
    if( rs.method )
    {
        if( evc0 )
            HIPCHECK( hipEventRecord( evc0, st ) );
        ddi_rs_launch_field( rs, spins_soa, h_soa, st );
        if( evc1 )
            HIPCHECK( hipEventRecord( evc1, st ) );
        return;
    }
    ddi_launch_field( d, spins_soa, h_soa, nos, st, evc0, evc1 );
}

// ---------------------------------------------------------------------------
// Hamiltonian upload
// ---------------------------------------------------------------------------
FastDiv make_fast_div( int d )
{
    if( d < 1 )
        throw std::runtime_error( "lattice dimension must be positive" );
    int L = 0;
    while( ( 1LL << L ) < d )
        ++L;
    const int S             = 31 + L;
    const unsigned long long m = ( ( 1ULL << S ) / (unsigned long long)d ) + 1;
    FastDiv f;
    f.m_lo  = (uint32_t)( m & 0xffffffffULL );
    f.m_hi  = (uint32_t)( m >> 32 );
    f.shift = S;
    f.d     = d;
    return f;
}

void HamUpload::upload( const Geometry & g, const Hamiltonian & h )
{
    HamDev H{};
    std::vector<PairDev> pv = build_ham_dev( g, h, H );
    pairs.resize( sizeof( PairDev ) * std::max<size_t>( 1, pv.size() ) );
    if( !pv.empty() )
        HIPCHECK( hipMemcpy( pairs.as<void>(), pv.data(), sizeof( PairDev ) * pv.size(), hipMemcpyHostToDevice ) );
    H.exch = pairs.as<PairDev>();
    H.dmi  = pairs.as<PairDev>() + H.n_exch;
    if( !h.quads.empty() )
    {
        std::vector<QuadDev> qv;
        for( size_t k = 0; k < h.quads.size(); ++k )
        {
            const Quadruplet & q = h.quads[k];
            if( q.i < 0 || q.i >= g.nb || q.j < 0 || q.j >= g.nb || q.k < 0 || q.k >= g.nb || q.l < 0 || q.l >= g.nb )
                throw std::runtime_error( "quadruplet with a basis index outside the cell" );
            QuadDev d{};
            d.i = q.i, d.j = q.j, d.k = q.k, d.l = q.l;
            for( int c = 0; c < 3; ++c )
            {
                d.dj[c] = q.dj[c];
                d.dk[c] = q.dk[c];
                d.dl[c] = q.dl[c];
            }
            d.Q = h.quad_Q[k];
            qv.push_back( d );
        }
        quads.resize( sizeof( QuadDev ) * qv.size() );
        HIPCHECK( hipMemcpy( quads.as<void>(), qv.data(), sizeof( QuadDev ) * qv.size(), hipMemcpyHostToDevice ) );
        H.n_quad       = (int)qv.size();
        H.quad_in_step = h.quadruplets_in_gradient_and_energy() ? 1 : 0;
        H.quad         = quads.as<QuadDev>();
    }
    if( H.has_site_K )
    {
        site_K.resize( sizeof( double ) * g.nos );
        HIPCHECK( hipMemcpy( site_K.as<void>(), h.site_K.data(), sizeof( double ) * g.nos, hipMemcpyHostToDevice ) );
        H.site_K = site_K.as<double>();
    }
    dev         = H;
    version     = h.version;
    ddi_version = h.ddi_version;
}

std::vector<PairDev> build_ham_dev( const Geometry & g, const Hamiltonian & h, HamDev & H )
{
    std::vector<PairDev> pv( h.exch_pairs.size() + h.dmi_pairs.size() );
    fill_ham_dev( g, h, H, pv.data(), (int)pv.size() );
    return pv;
}

int fill_ham_dev( const Geometry & g, const Hamiltonian & h, HamDev & H, PairDev * pv, int capacity )
{
    const size_t n_pairs = h.exch_pairs.size() + h.dmi_pairs.size();
    if( n_pairs > (size_t)capacity )
        throw std::runtime_error( "pair table larger than its buffer (" + std::to_string( capacity ) + " pairs)" );
    if( g.nb > MAX_BASIS )
        throw std::runtime_error( "too many basis atoms (max " + std::to_string( MAX_BASIS ) + ")" );
    if( (int)h.aniso_idx.size() > MAX_BASIS || (int)h.cubic_idx.size() > MAX_BASIS )
        throw std::runtime_error( "too many anisotropy entries" );
    H     = HamDev{};
    H.Na  = g.n_cells[0];
    H.Nb  = g.n_cells[1];
    H.Nc  = g.n_cells[2];
    H.nb  = g.nb;
    H.nos = g.nos;
    H.div_nb = make_fast_div( g.nb );
    H.div_Na = make_fast_div( g.n_cells[0] );
    H.div_Nb = make_fast_div( g.n_cells[1] );
    for( int c = 0; c < 3; ++c )
        H.bc[c] = h.bc[c];
    H.has_zeeman = h.has_zeeman() ? 1 : 0;
    H.B          = h.field_magnitude;
    for( int c = 0; c < 3; ++c )
        H.Bn[c] = h.field_normal[c];
    for( int b = 0; b < g.nb; ++b )
    {
        H.mu[b]     = g.mu_basis[b];
        H.inv_mu[b] = 1.0 / g.mu_basis[b];
    }
    H.n_aniso = (int)h.aniso_idx.size();
    for( int k = 0; k < H.n_aniso; ++k )
    {
        H.aniso_idx[k] = h.aniso_idx[k];
        H.aniso_K[k]   = h.aniso_K[k];
        for( int c = 0; c < 3; ++c )
            H.aniso_n[k][c] = h.aniso_n[k][c];
    }
    H.n_cubic = (int)h.cubic_idx.size();
    for( int k = 0; k < H.n_cubic; ++k )
    {
        H.cubic_idx[k] = h.cubic_idx[k];
        H.cubic_K[k]   = h.cubic_K[k];
    }
    size_t np = 0;
    for( size_t k = 0; k < h.exch_pairs.size(); ++k )
    {
        const Pair & p = h.exch_pairs[k];
        PairDev q{};
        q.i = p.i, q.j = p.j, q.da = p.t[0], q.db = p.t[1], q.dc = p.t[2];
        q.J   = h.exch_J[k];
        q.pad = k < h.exch_shell.size() ? h.exch_shell[k] : 0;
        pv[np++] = q;
    }
    for( size_t k = 0; k < h.dmi_pairs.size(); ++k )
    {
        const Pair & p = h.dmi_pairs[k];
        PairDev q{};
        q.i = p.i, q.j = p.j, q.da = p.t[0], q.db = p.t[1], q.dc = p.t[2];
        q.D   = h.dmi_D[k];
        q.pad = k < h.dmi_shell.size() ? h.dmi_shell[k] : 0;
        for( int c = 0; c < 3; ++c )
            q.n[c] = h.dmi_n[k][c];
        pv[np++] = q;
    }
    for( size_t k = 0; k < np; ++k )
    {
        PairDev & q = pv[k];
        if( q.i < 0 || q.i >= g.nb || q.j < 0 || q.j >= g.nb )
            throw std::runtime_error( "pair references a basis atom that does not exist" );
        const long long d = ( q.j - q.i ) + (long long)g.nb * ( q.da + (long long)g.n_cells[0]
                                                                        * ( q.db + (long long)g.n_cells[1] * q.dc ) );
        q.delta = (int)d; // only used where the site is at least |translation| away from every boundary
    }
    H.n_exch = (int)h.exch_pairs.size();
    H.n_dmi  = (int)h.dmi_pairs.size();
    for( int k = 0; k < H.n_exch; ++k )
    {
        H.halo[0] = std::max( H.halo[0], std::abs( pv[k].da ) );
        H.halo[1] = std::max( H.halo[1], std::abs( pv[k].db ) );
        H.halo[2] = std::max( H.halo[2], std::abs( pv[k].dc ) );
    }
    // stage-kernel row tile: 16 rows by default (256x256x128 Depondt+DDI: 0.308 / 0.300 /
    // 0.287 / 0.286 ms per stage at 4 / 8 / 16 / 32 rows); SPIRIT_HIP_STAGE_ROW_TILE (a power
    // of two, 0 for lattice order) overrides it for launch-shape sweeps
    H.row_tile_log2 = 4;
    if( const char * rt = std::getenv( "SPIRIT_HIP_STAGE_ROW_TILE" ) )
    {
        const int v = std::atoi( rt );
        int l       = 0;
        while( ( 1 << l ) < v && l < 12 )
            ++l;
        H.row_tile_log2 = ( v > 0 && ( 1 << l ) == v ) ? l : 0;
    }
    for( int k = 0; k < GATHER_PAIRS && k < H.n_exch; ++k )
        H.exch_g[k] = pv[k];
    for( int k = 0; k < GATHER_PAIRS && k < (int)h.dmi_pairs.size(); ++k )
        H.dmi_g[k] = pv[H.n_exch + k];
    H.dmi_same = H.n_dmi <= H.n_exch ? 1 : 0;
    for( int k = 0; k < H.n_dmi && H.dmi_same; ++k )
    {
        const PairDev &e = pv[k], &m = pv[H.n_exch + k];
        if( e.i != m.i || e.j != m.j || e.da != m.da || e.db != m.db || e.dc != m.dc )
            H.dmi_same = 0;
    }
    H.exch = nullptr;
    H.dmi  = nullptr;
    if( !h.site_K.empty() )
    {
        if( (int)h.site_K.size() != g.nos )
            throw std::runtime_error( "per-site anisotropy has the wrong length" );
        H.has_site_K = 1;
        H.site_K     = nullptr;
        for( int c = 0; c < 3; ++c )
            H.site_n[c] = h.site_n[c];
    }
    H.has_ddi = ( h.ddi_method == DDI_FFT || h.ddi_method == DDI_CUTOFF ) ? 1 : 0;
    return (int)np;
}

// ---------------------------------------------------------------------------
// LLG runner
// ---------------------------------------------------------------------------
void aos_to_soa( const double * aos, double * soa, long long n )
{
    for( long long i = 0; i < n; ++i )
    {
        soa[i]         = aos[3 * i];
        soa[n + i]     = aos[3 * i + 1];
        soa[2 * n + i] = aos[3 * i + 2];
    }
}
void soa_to_aos( const double * soa, double * aos, long long n )
{
    for( long long i = 0; i < n; ++i )
    {
        aos[3 * i]     = soa[i];
        aos[3 * i + 1] = soa[n + i];
        aos[3 * i + 2] = soa[2 * n + i];
    }
}

static LLGDev make_llg_dev( const LLGParameters & p, const Geometry & g )
{
    LLGDev L{};
    L.dt                  = p.dt;
    L.damping             = p.damping;
    L.temperature         = p.temperature;
    L.direct_minimization = p.direct_minimization ? 1 : 0;
    // Prepare_Thermal_Field draws a field at T > 0 or with a temperature gradient (Method_LLG.cpp:72)
    L.thermal             = ( p.temperature > 0 || p.temperature_gradient_inclination != 0 ) ? 1 : 0;
    L.seed                = p.seed;
    L.force_convergence   = p.force_convergence;
    L.vf_dtg              = p.dt * gamma_e / mu_B / ( 1 + p.damping * p.damping );
    // spin-transfer torque, Method_LLG.cpp:150-157, 182-212 (applied for a_j > 0 only)
    L.stt = STT_OFF;
    if( p.stt_magnitude > 0 )
    {
        const double dtg = L.vf_dtg, aj = p.stt_magnitude, al = p.damping, be = p.beta;
        const Vec3 & j   = p.stt_polarisation_normal;
        if( p.stt_use_gradient )
        {
            L.stt    = STT_GRADIENT;
            L.stt_c1 = dtg * aj * ( al - be );
            L.stt_c2 = dtg * aj * ( 1 + be * al );
            // Vectormath::jacobian's J = M B^-1 (Vectormath.cpp:827-838, 896) applied to j is
            // M (B^-1 j): the differences along the three translations weighted by w = B^-1 j
            double B[3][3];
            for( int r = 0; r < 3; ++r )
                for( int k = 0; k < 3; ++k )
                    B[r][k] = g.lattice_constant * g.bravais[k][r];
            const double det = B[0][0] * ( B[1][1] * B[2][2] - B[1][2] * B[2][1] )
                               - B[0][1] * ( B[1][0] * B[2][2] - B[1][2] * B[2][0] )
                               + B[0][2] * ( B[1][0] * B[2][1] - B[1][1] * B[2][0] );
            if( det == 0 )
                throw std::runtime_error( "STT gradient: the Bravais vectors are linearly dependent" );
            double Bi[3][3];
            for( int r = 0; r < 3; ++r )
                for( int k = 0; k < 3; ++k )
                {
                    // adjugate: cofactor of B[k][r]
                    const int r1 = ( k + 1 ) % 3, r2 = ( k + 2 ) % 3, c1 = ( r + 1 ) % 3, c2 = ( r + 2 ) % 3;
                    Bi[r][k] = ( B[r1][c1] * B[r2][c2] - B[r1][c2] * B[r2][c1] ) / det;
                }
            for( int t = 0; t < 3; ++t )
                L.stt_w[t] = Bi[t][0] * j[0] + Bi[t][1] * j[1] + Bi[t][2] * j[2];
        }
        else
        {
            L.stt    = STT_MONOLAYER;
            L.stt_c1 = -dtg * aj * ( al - be );
            L.stt_c2 = -dtg * aj * ( 1 + be * al );
            for( int q = 0; q < 3; ++q )
                L.stt_p[q] = j[q];
        }
    }
    L.temp_site = nullptr;
    return L;
}

// Vectormath::get_gradient_distribution (Vectormath.cpp:633-652): T_i = incl (d . r_i) +
// (T - incl min(d . bounds_min, d . bounds_max)), cut to [0, 1e30]
static std::vector<double> temperature_distribution( const LLGParameters & p, const Geometry & g )
{
    Vec3 d         = p.temperature_gradient_direction;
    const double n = std::sqrt( d[0] * d[0] + d[1] * d[1] + d[2] * d[2] );
    if( n > 0 )
        for( int q = 0; q < 3; ++q )
            d[q] /= n;
    const double incl  = p.temperature_gradient_inclination;
    const double bmin  = g.bounds_min[0] * d[0] + g.bounds_min[1] * d[1] + g.bounds_min[2] * d[2];
    const double bmax  = g.bounds_max[0] * d[0] + g.bounds_max[1] * d[1] + g.bounds_max[2] * d[2];
    const double shift = p.temperature - incl * std::min( bmin, bmax );
    std::vector<double> T( g.nos );
    for( int i = 0; i < g.nos; ++i )
    {
        const double * r = &g.positions[3 * (size_t)i];
        double t         = incl * ( d[0] * r[0] + d[1] * r[1] + d[2] * r[2] );
        t += shift;
        T[i] = std::min( std::max( 0.0, t ), 1e30 );
    }
    return T;
}

void LLGRunner::apply_llg( size_t k )
{
    const LLGParameters & p = specs[k].llg;
    imgs_host[k].llg        = make_llg_dev( p, *specs[k].geometry );
    if( b_temp.size() < specs.size() )
        b_temp.resize( specs.size() );
    if( p.temperature_gradient_inclination != 0 )
    {
        const std::vector<double> T = temperature_distribution( p, *specs[k].geometry );
        if( !b_temp[k] )
            b_temp[k].reset( new DeviceBuffer( sizeof( double ) * T.size() ) );
        HIPCHECK( hipMemcpy( b_temp[k]->as<void>(), T.data(), sizeof( double ) * T.size(), hipMemcpyHostToDevice ) );
        imgs_host[k].llg.temp_site = b_temp[k]->as<double>();
    }
    int any = 0;
    for( auto & I : imgs_host )
        any = any || ( I.llg.stt == STT_GRADIENT && !I.llg.direct_minimization );
    if( any != stt_gradient_any )
        release_step_graph(); // the captured stage kernels are the other instantiation
    stt_gradient_any = any;
}

LLGRunner::LLGRunner( const std::vector<ImageSpec> & images, int solver_ ) : specs( images ), solver( solver_ )
{
    if( solver != SOLVER_DEPONDT && solver != SOLVER_HEUN && solver != SOLVER_SIB && solver != SOLVER_RK4
        && !min_is_solver( solver ) )
        throw std::runtime_error( "LLG solver " + std::to_string( solver ) + " is not available in the HIP engine" );
    if( specs.empty() )
        throw std::runtime_error( "LLGRunner: no images" );
    if( min_is_solver( solver ) && specs.size() != 1 )
        throw std::runtime_error( "the direct-minimisation solvers run one image at a time" );
    HIPCHECK( hipStreamCreateWithFlags( &stream, hipStreamNonBlocking ) );
    long long off = 0;
    n_blocks      = 0;
    for( size_t k = 0; k < specs.size(); ++k )
    {
        auto & sp = specs[k];
        hams.emplace_back( new HamUpload() );
        hams.back()->upload( *sp.geometry, *sp.hamiltonian );
        ImageDev I{};
        I.ham         = hams.back()->dev;
        I.llg         = make_llg_dev( sp.llg, *sp.geometry );
        I.off         = off;
        I.block_begin = n_blocks;
        I.n_blocks    = ( sp.geometry->nos + LLG_THREADS - 1 ) / LLG_THREADS;
        for( int b = 0; b < I.n_blocks; ++b )
            blk_img_host.push_back( (int)k );
        n_blocks += I.n_blocks;
        off += sp.geometry->nos;
        imgs_host.push_back( I );
        if( sp.hamiltonian->ddi_method == DDI_FFT || sp.hamiltonian->ddi_method == DDI_CUTOFF )
        {
            ddi.emplace_back( new DdiPlan() );
            ddi.back()->build( *sp.geometry, *sp.hamiltonian, stream );
            any_ddi = true;
        }
        else
        {
            // FMM: Gradient_DDI (Hamiltonian_Heisenberg.cpp:866-878) evaluates no field for it
            ddi.emplace_back( nullptr );
        }
    }
    n_total         = off;
    for( size_t k = 0; k < specs.size(); ++k )
        apply_llg( k );
    const size_t f3 = sizeof( double ) * 3 * (size_t)n_total;
    b_s.resize( f3 );
    b_sp.resize( f3 );
    if( solver == SOLVER_RK4 )
    {
        b_sp2.resize( f3 );
        b_aux2.resize( f3 );
    }
    b_aux.resize( f3 );
    b_eff.resize( f3 );
    b_efft.resize( f3 );
    if( any_ddi )
    {
        b_h.resize( f3 );
        HIPCHECK( hipMemset( b_h.as<void>(), 0, f3 ) );
    }
    b_pE.resize( sizeof( double ) * n_blocks );
    b_pT.resize( sizeof( double ) * n_blocks );
    alloc_thermal_field();
    b_stop.resize( sizeof( int ) );
    b_stopstep.resize( sizeof( long long ) * specs.size() );
    b_img_stop.resize( sizeof( int ) * specs.size() );
    b_nstop.resize( sizeof( int ) );
    b_step.resize( sizeof( uint64_t ) );
    HIPCHECK( hipMemset( b_img_stop.as<void>(), 0, sizeof( int ) * specs.size() ) );
    HIPCHECK( hipMemset( b_nstop.as<void>(), 0, sizeof( int ) ) );
    HIPCHECK( hipMemset( b_stop.as<void>(), 0, sizeof( int ) ) );
    img_stopped.assign( specs.size(), 0 );
    steps_last.assign( specs.size(), 0 );
    b_obs.resize( sizeof( double ) * 2 * specs.size() );
    HIPCHECK( hipMemset( b_obs.as<void>(), 0, sizeof( double ) * 2 * specs.size() ) );
    HIPCHECK( hipMemset( b_eff.as<void>(), 0, f3 ) );
    obs_host.assign( 2 * specs.size(), 0.0 );
    b_blk.resize( sizeof( int ) * blk_img_host.size() );
    HIPCHECK( hipMemcpy( b_blk.as<void>(), blk_img_host.data(), sizeof( int ) * blk_img_host.size(),
                         hipMemcpyHostToDevice ) );
    upload_images();
    upload_spins();
    if( min_is_solver( solver ) )
        setup_minimiser();
    else
        tune_stage_tiles();
}


// Buffers and scalar state of a direct-minimisation run, zero as after the solver's
// Initialize (Solver_VP.hpp:2-18, Solver_VP_OSO.hpp:2-19, Solver_LBFGS_OSO.hpp:12-31,
// Solver_LBFGS_Atlas.hpp:13-42); the atlas charts from the signs of the initial s_z.
void LLGRunner::setup_minimiser()
{
    const size_t nos = (size_t)n_total;
    const int dim    = solver == SOLVER_LBFGS_ATLAS ? 2 : 3;
    auto zeroed      = [&]( DeviceBuffer & b, size_t bytes ) {
        b.resize( bytes );
        HIPCHECK( hipMemset( b.as<void>(), 0, bytes ) );
    };
    const size_t f3 = sizeof( double ) * 3 * nos, fd = sizeof( double ) * dim * nos;
    zeroed( b_mv, f3 );
    zeroed( b_mprev, f3 );
    zeroed( b_mf, f3 );
    zeroed( b_mgrad, fd );
    zeroed( b_mgpr, fd );
    zeroed( b_msd, fd );
    zeroed( b_mq, fd );
    for( int k = 0; k < MIN_MEMORY; ++k )
    {
        zeroed( b_mda[k], fd );
        zeroed( b_mdg[k], fd );
    }
    zeroed( b_mpart, sizeof( double ) * MIN_SLOTS * (size_t)n_blocks );
    zeroed( b_msc, sizeof( MinScalars ) );
    b_ma3.resize( sizeof( double ) * nos );
    std::vector<double> a3( nos, 1.0 );
    for( size_t i = 0; i < nos; ++i )
        a3[i] = specs[0].spins_aos[3 * i + 2] > 0 ? 1.0 : -1.0;
    HIPCHECK( hipMemcpy( b_ma3.as<void>(), a3.data(), sizeof( double ) * nos, hipMemcpyHostToDevice ) );
    MinArgs & M = min_args;
    M.v         = b_mv.as<double>();
    M.prev      = b_mprev.as<double>();
    M.f         = b_mf.as<double>();
    M.grad      = b_mgrad.as<double>();
    M.grad_pr   = b_mgpr.as<double>();
    M.sd        = b_msd.as<double>();
    M.q         = b_mq.as<double>();
    for( int k = 0; k < MIN_MEMORY; ++k )
    {
        M.da[k] = b_mda[k].as<double>();
        M.dg[k] = b_mdg[k].as<double>();
    }
    M.a3       = b_ma3.as<double>();
    M.part     = b_mpart.as<double>();
    M.F_ext    = nullptr;
    M.sc       = b_msc.as<MinScalars>();
    M.n_blocks = n_blocks;
    M.noi      = 1;
    M.bpi      = n_blocks;
    M.dim      = dim;
    M.maxmove  = solver == SOLVER_LBFGS_ATLAS ? 0.05 : Pi / 200.0; // Solver_LBFGS_*.hpp Initialize
}

LLGRunner::~LLGRunner()
{
    if( stream )
    {
        (void)hipStreamSynchronize( stream );
        release_step_graph();
        for( auto e : ev_pool )
            (void)hipEventDestroy( e );
        (void)hipStreamDestroy( stream );
    }
}

void LLGRunner::set_profiling( bool on )
{
    profiling = on;
}

hipEvent_t LLGRunner::next_event()
{
    if( ev_used == ev_pool.size() )
    {
        hipEvent_t e;
        HIPCHECK( hipEventCreate( &e ) );
        ev_pool.push_back( e );
    }
    return ev_pool[ev_used++];
}

void LLGRunner::collect_profile()
{
    for( auto & p : pending )
    {
        float ms = 0;
        HIPCHECK( hipEventElapsedTime( &ms, p.second.first, p.second.second ) );
        bool found = false;
        for( auto & e : prof )
            if( e.first == p.first )
            {
                e.second.ms += ms;
                e.second.count += 1;
                found = true;
            }
        if( !found )
        {
            ProfileEntry pe;
            pe.ms    = ms;
            pe.count = 1;
            prof.push_back( { p.first, pe } );
        }
    }
    pending.clear();
    ev_used = 0;
}

LLGRunner::ProfileEntry LLGRunner::profile( const std::string & name ) const
{
    for( auto & e : prof )
        if( e.first == name )
            return e.second;
    return ProfileEntry{};
}

void LLGRunner::reset_profile()
{
    prof.clear();
}

void LLGRunner::upload_images()
{
    b_imgs.resize( sizeof( ImageDev ) * imgs_host.size() );
    HIPCHECK( hipMemcpy( b_imgs.as<void>(), imgs_host.data(), sizeof( ImageDev ) * imgs_host.size(),
                         hipMemcpyHostToDevice ) );
}

void LLGRunner::set_llg( int img, const LLGParameters & p )
{
    specs[img].llg = p;
    apply_llg( img );
    upload_images();
    alloc_thermal_field();
}

// the step's stored thermal field, once any image runs at T > 0
void LLGRunner::alloc_thermal_field()
{
    bool thermal = false;
    for( auto & I : imgs_host )
        thermal = thermal || I.llg.thermal;
    if( thermal && b_xi.size() == 0 )
    {
        release_step_graph();
        b_xi.resize( sizeof( double ) * 3 * (size_t)n_total );
    }
}

void LLGRunner::refresh_hamiltonian()
{
    release_step_graph(); // DDI plans and buffers may be rebuilt
    for( size_t k = 0; k < specs.size(); ++k )
    {
        const uint64_t old_ddi = hams[k]->ddi_version;
        hams[k]->upload( *specs[k].geometry, *specs[k].hamiltonian );
        imgs_host[k].ham = hams[k]->dev;
        if( tuned_row_tile >= 0 )
            imgs_host[k].ham.row_tile_log2 = tuned_row_tile;
        if( specs[k].hamiltonian->ddi_method == DDI_FFT || specs[k].hamiltonian->ddi_method == DDI_CUTOFF )
        {
            if( !ddi[k] || old_ddi != specs[k].hamiltonian->ddi_version )
            {
                ddi[k].reset( new DdiPlan() );
                ddi[k]->build( *specs[k].geometry, *specs[k].hamiltonian, stream );
            }
            if( !any_ddi )
            {
                any_ddi = true;
                b_h.resize( sizeof( double ) * 3 * (size_t)n_total );
                HIPCHECK( hipMemset( b_h.as<void>(), 0, b_h.size() ) );
            }
        }
        else
            ddi[k].reset();
    }
    upload_images();
}

// Transfers of whole batches go through page-locked host memory: the runtime copies
// to and from pageable memory in pieces through its own staging (about 5 GB/s here), and
// page-locked memory moves at the link's rate.  One buffer per process, grown on demand
// and kept to the end (page-locking costs about as much as one pageable transfer of the
// same size); a transfer holds its lock from the first copy into it to the last copy out.
// Below PINNED_MIN_BYTES the images' mirrors are copied directly.
static constexpr size_t PINNED_MIN_BYTES = size_t( 8 ) << 20;
static std::mutex pinned_mutex;
static double * pinned_buf    = nullptr;
static size_t pinned_capacity = 0; // bytes

static double * pinned_reserve( size_t bytes )
{
    if( bytes > pinned_capacity )
    {
        if( pinned_buf )
            HIPCHECK( hipHostFree( pinned_buf ) );
        pinned_buf      = nullptr;
        pinned_capacity = 0;
        HIPCHECK( hipHostMalloc( (void **)&pinned_buf, bytes, hipHostMallocDefault ) );
        pinned_capacity = bytes;
    }
    return pinned_buf;
}

struct HostCopy
{
    void * dst;
    const void * src;
    size_t bytes;
};

// Host copies on up to 8 threads, in pieces of at most 4 MiB taken in turn (one thread
// copies about 10 GB/s; the batch's 256 mirrors would otherwise cost more than the link).
static void parallel_host_copies( const std::vector<HostCopy> & copies )
{
    constexpr size_t PIECE = size_t( 4 ) << 20;
    std::vector<HostCopy> pieces;
    size_t total = 0;
    for( const auto & c : copies )
    {
        total += c.bytes;
        for( size_t o = 0; o < c.bytes; o += PIECE )
            pieces.push_back( { (char *)c.dst + o, (const char *)c.src + o, std::min( PIECE, c.bytes - o ) } );
    }
    const unsigned hw = std::max( 1u, std::thread::hardware_concurrency() );
    const int n_thr   = (int)std::min<size_t>( { (size_t)8, (size_t)hw, pieces.size() } );
    std::atomic<size_t> next( 0 );
    auto work = [&]() {
        for( size_t k = next++; k < pieces.size(); k = next++ )
            std::memcpy( pieces[k].dst, pieces[k].src, pieces[k].bytes );
    };
    if( n_thr <= 1 || total < PINNED_MIN_BYTES )
    {
        work();
        return;
    }
    std::vector<std::thread> pool;
    for( int t = 1; t < n_thr; ++t )
    {
        try
        {
            pool.emplace_back( work );
        }
        catch( ... )
        {
            break; // fewer threads: the pieces are taken from the shared counter either way
        }
    }
    work();
    for( auto & t : pool )
        t.join();
}

// The host mirrors are AoS, the device fields SoA: each image's mirror is copied as it is into
// a device staging buffer and transposed there (a host transpose of the whole batch costs
// more than the copies).
void LLGRunner::upload_spins()
{
    const size_t bytes = sizeof( double ) * 3 * (size_t)n_total;
    if( b_stage.size() < bytes )
        b_stage.resize( bytes );
    double * stage = b_stage.as<double>();
    HIPCHECK( hipStreamSynchronize( stream ) ); // no kernel of an earlier call still reads the stage
    if( bytes >= PINNED_MIN_BYTES )
    {
        std::lock_guard<std::mutex> guard( pinned_mutex );
        double * pin = pinned_reserve( bytes );
        std::vector<HostCopy> copies;
        for( size_t k = 0; k < specs.size(); ++k )
            copies.push_back( { pin + 3 * imgs_host[k].off, specs[k].spins_aos,
                                sizeof( double ) * 3 * (size_t)specs[k].geometry->nos } );
        parallel_host_copies( copies );
        HIPCHECK( hipMemcpyAsync( stage, pin, bytes, hipMemcpyHostToDevice, stream ) );
        for( size_t k = 0; k < specs.size(); ++k )
        {
            const long long base = 3 * imgs_host[k].off;
            device_aos_to_soa( stage + base, b_s.as<double>() + base, specs[k].geometry->nos, stream );
        }
        HIPCHECK( hipGetLastError() );
        HIPCHECK( hipStreamSynchronize( stream ) );
        return;
    }
    for( size_t k = 0; k < specs.size(); ++k )
    {
        const long long base = 3 * imgs_host[k].off;
        const int nos        = specs[k].geometry->nos;
        HIPCHECK( hipMemcpyAsync( stage + base, specs[k].spins_aos, sizeof( double ) * 3 * (size_t)nos,
                                  hipMemcpyHostToDevice, stream ) );
        device_aos_to_soa( stage + base, b_s.as<double>() + base, nos, stream );
    }
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipStreamSynchronize( stream ) );
}

// Copies the AoS device stage of the images with a destination (nullptr: skipped) to their
// host arrays, with one synchronisation.  Runs of consecutive images move in one copy.
void LLGRunner::stage_to_host( const std::vector<double *> & dst )
{
    const double * stage = b_stage.as<double>();
    size_t bytes         = 0;
    for( size_t k = 0; k < specs.size() && k < dst.size(); ++k )
        if( dst[k] )
            bytes += sizeof( double ) * 3 * (size_t)specs[k].geometry->nos;
    if( bytes < PINNED_MIN_BYTES )
    {
        for( size_t k = 0; k < specs.size() && k < dst.size(); ++k )
            if( dst[k] )
                HIPCHECK( hipMemcpyAsync( dst[k], stage + 3 * imgs_host[k].off,
                                          sizeof( double ) * 3 * (size_t)specs[k].geometry->nos,
                                          hipMemcpyDeviceToHost, stream ) );
        HIPCHECK( hipStreamSynchronize( stream ) );
        return;
    }
    std::lock_guard<std::mutex> guard( pinned_mutex );
    double * pin = pinned_reserve( sizeof( double ) * 3 * (size_t)n_total );
    std::vector<HostCopy> copies;
    for( size_t k = 0; k < specs.size() && k < dst.size(); )
    {
        if( !dst[k] )
        {
            ++k;
            continue;
        }
        size_t e = k;
        while( e < specs.size() && e < dst.size() && dst[e] )
            ++e;
        const long long b0 = 3 * imgs_host[k].off;
        const long long b1 = 3 * ( imgs_host[e - 1].off + specs[e - 1].geometry->nos );
        HIPCHECK( hipMemcpyAsync( pin + b0, stage + b0, sizeof( double ) * (size_t)( b1 - b0 ), hipMemcpyDeviceToHost,
                                  stream ) );
        for( ; k < e; ++k )
            copies.push_back( { dst[k], pin + 3 * imgs_host[k].off,
                                sizeof( double ) * 3 * (size_t)specs[k].geometry->nos } );
    }
    HIPCHECK( hipStreamSynchronize( stream ) );
    parallel_host_copies( copies );
}

void LLGRunner::download_spins()
{
    if( b_stage.size() < sizeof( double ) * 3 * (size_t)n_total )
        b_stage.resize( sizeof( double ) * 3 * (size_t)n_total );
    double * stage = b_stage.as<double>();
    std::vector<double *> dst( specs.size() );
    for( size_t k = 0; k < specs.size(); ++k )
    {
        const long long base = 3 * imgs_host[k].off;
        device_soa_to_aos( b_s.as<double>() + base, stage + base, specs[k].geometry->nos, stream );
        dst[k] = specs[k].spins_aos;
    }
    HIPCHECK( hipGetLastError() );
    stage_to_host( dst );
}

void LLGRunner::download_effective_field( int img, double * aos )
{
    std::vector<double *> dst( specs.size(), nullptr );
    dst[img] = aos;
    download_effective_fields( dst );
}

// The effective fields of the images with a destination (AoS host arrays; nullptr: skipped),
// projected, transposed to AoS on the device and copied, with one synchronisation.
void LLGRunner::download_effective_fields( const std::vector<double *> & aos )
{
    if( b_stage.size() < sizeof( double ) * 3 * (size_t)n_total )
        b_stage.resize( sizeof( double ) * 3 * (size_t)n_total );
    double * stage = b_stage.as<double>();
    for( size_t k = 0; k < specs.size() && k < aos.size(); ++k )
    {
        if( !aos[k] )
            continue;
        const int nos        = specs[k].geometry->nos;
        const long long base = 3 * imgs_host[k].off;
        const double * src   = b_eff.as<double>() + base;
        if( eff_raw )
        {
            // P_perp(s) F1 with the spins of the last observed step (every run ends on one)
            llg_launch_project_field( src, b_s.as<double>() + base, b_efft.as<double>() + base, nos, stream );
            src = b_efft.as<double>() + base;
        }
        device_soa_to_aos( src, stage + base, nos, stream );
    }
    HIPCHECK( hipGetLastError() );
    stage_to_host( aos );
}

void LLGRunner::set_injected_noise( const double * noise, long n_steps )
{
    if( !noise || n_steps <= 0 )
    {
        b_noise.free();
        noise_steps = noise_pos = 0;
        return;
    }
    const size_t n = (size_t)n_steps * 3 * (size_t)n_total;
    b_noise.resize( sizeof( double ) * n );
    HIPCHECK( hipMemcpy( b_noise.as<void>(), noise, sizeof( double ) * n, hipMemcpyHostToDevice ) );
    noise_steps = n_steps;
    noise_pos   = 0;
}

void LLGRunner::ddi_field( const double * x, bool prof_events )
{
    for( size_t k = 0; k < specs.size(); ++k )
        if( ddi[k] )
        {
            const long long base = 3 * imgs_host[k].off;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if( prof_events )
            {
                e0 = next_event();
                e1 = next_event();
                pending.push_back( { "ddi_conv", { e0, e1 } } );
            }
            ddi[k]->field( x + base, b_h.as<double>() + base, specs[k].geometry->nos, stream, e0, e1 );
        }
}

static StepArgs base_args_of( DeviceBuffer & imgs, DeviceBuffer & blk, DeviceBuffer & s, DeviceBuffer & sp,
                              DeviceBuffer & sp2, DeviceBuffer & aux, DeviceBuffer & aux2, DeviceBuffer & h,
                              DeviceBuffer & eff, DeviceBuffer & efft, DeviceBuffer & pE, DeviceBuffer & pT,
                              DeviceBuffer & stop, DeviceBuffer & img_stop, DeviceBuffer & xi )
{
    StepArgs A{};
    A.imgs    = imgs.as<ImageDev>();
    A.blk_img = blk.as<int>();
    A.s       = s.as<double>();
    A.sp      = sp.as<double>();
    A.sp2     = sp2.as<double>();
    A.aux     = aux.as<double>();
    A.aux2    = aux2.as<double>();
    A.h       = h.as<double>();
    A.eff     = eff.as<double>();
    (void)efft;
    A.part_E  = pE.as<double>();
    A.part_T  = pT.as<double>();
    A.stop     = stop.as<int>();
    A.img_stop = img_stop.as<int>();
    A.xi       = xi.as<double>();
    return A;
}

static StepArgs with_stt( StepArgs A, int stt_gradient )
{
    A.stt_gradient = stt_gradient;
    return A;
}

#define STEP_ARGS()                                                                                                    \
    with_stt( base_args_of( b_imgs, b_blk, b_s, b_sp, b_sp2, b_aux, b_aux2, b_h, b_eff, b_efft, b_pE, b_pT, b_stop,   \
                            b_img_stop, b_xi ),                                                                        \
              stt_gradient_any )

// The stage kernels' row tile (rows along b swept through the c planes) by timing: the first
// solver stage at tiles of 8, 16 and 32 rows (64 measured 5-8% slower than 32), where every
// image's rows hold whole blocks and Nb is a multiple of the tile; the tile with the lowest
// median of seven timed launches is kept for the run (lattices below 2^20 spins keep the
// default).  SPIRIT_HIP_STAGE_ROW_TILE fixes the tile instead.
void LLGRunner::tune_stage_tiles()
{
    if( std::getenv( "SPIRIT_HIP_STAGE_ROW_TILE" ) || n_total < ( 1LL << 20 ) )
        return;
    std::vector<int> cand;
    for( int tl : { 3, 4, 5 } )
    {
        bool ok = true;
        for( auto & I : imgs_host )
        {
            const HamDev & H = I.ham;
            ok = ok && ( H.nb * H.Na ) % LLG_THREADS == 0 && H.Nb % ( 1 << tl ) == 0 && H.Nc > 1;
        }
        if( ok )
            cand.push_back( tl );
    }
    if( cand.size() < 2 )
        return;
    StepArgs A = STEP_ARGS();
    A.obs      = 0;
    hipEvent_t e0, e1;
    HIPCHECK( hipEventCreate( &e0 ) );
    HIPCHECK( hipEventCreate( &e1 ) );
    // rounds over the candidates in turn (a warm-up launch after each switch, then a timed
    // one), so that drifts of the clock hit every candidate alike; the median of each decides
    constexpr int ROUNDS = 7;
    std::vector<std::vector<double>> times( cand.size() );
    for( int round = 0; round < ROUNDS; ++round )
        for( size_t c = 0; c < cand.size(); ++c )
        {
            for( auto & I : imgs_host )
                I.ham.row_tile_log2 = cand[c];
            upload_images();
            llg_launch_stage( solver, 0, A, n_blocks, stream );
            HIPCHECK( hipEventRecord( e0, stream ) );
            llg_launch_stage( solver, 0, A, n_blocks, stream );
            HIPCHECK( hipEventRecord( e1, stream ) );
            HIPCHECK( hipEventSynchronize( e1 ) );
            float ms = 0;
            HIPCHECK( hipEventElapsedTime( &ms, e0, e1 ) );
            times[c].push_back( ms );
        }
    double best = 1e30;
    int pick    = cand[0];
    for( size_t c = 0; c < cand.size(); ++c )
    {
        std::sort( times[c].begin(), times[c].end() );
        const double t = times[c][ROUNDS / 2];
        char msg[80];
        std::snprintf( msg, sizeof( msg ), "stage row tile %d rows: %.4f ms (median)", 1 << cand[c], t );
        log_message( 6, msg );
        if( t < best )
        {
            best = t;
            pick = cand[c];
        }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    tuned_row_tile = pick;
    for( auto & I : imgs_host )
        I.ham.row_tile_log2 = pick;
    upload_images();
}

void LLGRunner::launch_step( const StepArgs & A )
{
    if( min_is_solver( solver ) )
    {
        if( any_ddi )
            ddi_field( A.s, profiling );
        min_launch_iteration( solver, A, min_args, (int)n_total, stream );
        return;
    }
    const int n_stages = solver == SOLVER_RK4 ? 4 : 2;
    for( int st = 0; st < n_stages; ++st )
    {
        const double * x = st == 0 ? A.s : ( ( solver == SOLVER_RK4 && st == 2 ) ? A.sp2 : A.sp );
        if( any_ddi )
            ddi_field( x, profiling );
        if( profiling )
        {
            hipEvent_t e0 = next_event(), e1 = next_event();
            HIPCHECK( hipEventRecord( e0, stream ) );
            llg_launch_stage( solver, st, A, n_blocks, stream );
            HIPCHECK( hipEventRecord( e1, stream ) );
            pending.push_back( { "llg_stage", { e0, e1 } } );
        }
        else
            llg_launch_stage( solver, st, A, n_blocks, stream );
    }
}

void LLGRunner::step( bool obs, bool count, bool check, double conv )
{
    (void)count;
    StepArgs A = STEP_ARGS();
    A.step     = dev_step ? 0 : step_counter; // with the device counter: offset 0 from it
    A.step_dev = dev_step ? b_step.as<uint64_t>() : nullptr;
    A.obs      = obs ? 1 : 0;
    A.noise    = nullptr;
    if( noise_steps > 0 )
    {
        if( noise_pos >= noise_steps )
            throw std::runtime_error( "injected thermal noise exhausted" );
        A.noise = b_noise.as<double>() + (size_t)noise_pos * 3 * (size_t)n_total;
        ++noise_pos;
    }
    launch_step( A );
    if( obs )
        llg_launch_finalize( A, (int)specs.size(), b_obs.as<double>(), b_img_stop.as<int>(), b_nstop.as<int>(),
                             b_stopstep.as<long long>(), (long long)A.step, conv, check ? 1 : 0, stream );
    if( dev_step )
        llg_launch_step_advance( b_step.as<uint64_t>(), 1, stream );
    ++step_counter;
}

// Long runs replay a captured graph of a few steps: the launches of the DDI passes, the
// solver stages and, on observed steps, the per-image reductions and convergence test, each
// step ending with the device step counter advanced.  The host then submits one graph instead
// of every launch; kernels and arguments are those of the launches one by one (the step index
// is a device base plus the step's offset in the graph, the base advanced once per replay), and
// stopped images and the run-wide stop flag are read on the device as before.
// SPIRIT_HIP_GRAPHS=0 switches graphs off.
bool LLGRunner::capture_step_graph( int len, int period, bool check, double conv )
{
    if( step_graph && graph_len == len && graph_period == period && graph_check == ( check ? 1 : 0 )
        && graph_conv == conv )
        return true;
    release_step_graph();
    StepArgs A = STEP_ARGS();
    A.step_dev = b_step.as<uint64_t>();
    A.noise    = nullptr;
    hipGraph_t graph = nullptr;
    if( hipStreamBeginCapture( stream, hipStreamCaptureModeThreadLocal ) != hipSuccess )
        return false;
    for( int k = 0; k < len; ++k )
    {
        // step k of the graph: index *b_step + k
        const bool obs = period > 0 && ( k + 1 ) % period == 0;
        A.obs          = obs ? 1 : 0;
        A.step         = (uint64_t)k;
        launch_step( A );
        if( obs )
            llg_launch_finalize( A, (int)specs.size(), b_obs.as<double>(), b_img_stop.as<int>(), b_nstop.as<int>(),
                                 b_stopstep.as<long long>(), (long long)k, conv, check ? 1 : 0, stream );
    }
    llg_launch_step_advance( b_step.as<uint64_t>(), len, stream );
    if( hipStreamEndCapture( stream, &graph ) != hipSuccess || !graph )
        return false;
    const hipError_t e = hipGraphInstantiate( &step_graph, graph, nullptr, nullptr, 0 );
    (void)hipGraphDestroy( graph );
    if( e != hipSuccess )
    {
        step_graph = nullptr;
        return false;
    }
    graph_len    = len;
    graph_period = period;
    graph_check  = check ? 1 : 0;
    graph_conv   = conv;
    return true;
}

void LLGRunner::release_step_graph()
{
    if( step_graph )
        (void)hipGraphExecDestroy( step_graph );
    step_graph   = nullptr;
    graph_len    = 0;
    graph_period = -1;
}

long LLGRunner::run( long n, int observe_every, bool check_convergence, double force_convergence )
{
    if( n <= 0 )
        return 0;
    // the host's view of which images have stopped; the device adds those that converge
    const int n_img = (int)specs.size();
    int n_stopped   = 0;
    for( int k = 0; k < n_img; ++k )
        n_stopped += img_stopped[k] ? 1 : 0;
    stop_host = n_stopped == n_img ? 1 : 0;
    nstop_host = n_stopped;
    HIPCHECK( hipMemcpyAsync( b_img_stop.as<void>(), img_stopped.data(), sizeof( int ) * n_img, hipMemcpyHostToDevice,
                              stream ) );
    HIPCHECK( hipMemcpyAsync( b_nstop.as<void>(), &nstop_host, sizeof( int ), hipMemcpyHostToDevice, stream ) );
    HIPCHECK( hipMemcpyAsync( b_stop.as<void>(), &stop_host, sizeof( int ), hipMemcpyHostToDevice, stream ) );
    const std::vector<int> was_stopped = img_stopped;
    const uint64_t first = step_counter;
    // Graph replays (not with injected noise, whose pointer changes every step, nor while
    // kernels are timed one by one).  With an observation period P <= GRAPH_MAX_STEPS a graph
    // holds whole periods and covers the steps of complete periods; otherwise (P = 0 or long
    // periods) a graph of unobserved steps covers the stretches between observed steps.
    const char * g_env  = std::getenv( "SPIRIT_HIP_GRAPHS" );
    const bool periodic = observe_every > 0 && observe_every <= GRAPH_MAX_STEPS;
    const int g_len     = periodic ? observe_every * ( ( GRAPH_STEPS + observe_every - 1 ) / observe_every )
                                   : GRAPH_STEPS;
    const bool graphs   = !( g_env && std::string( g_env ) == "0" ) && noise_steps == 0 && !profiling && n > g_len;
    dev_step = false;
    if( graphs )
    {
        HIPCHECK( hipMemcpyAsync( b_step.as<void>(), &step_counter, sizeof( uint64_t ), hipMemcpyHostToDevice,
                                  stream ) );
        HIPCHECK( hipStreamSynchronize( stream ) );
        if( capture_step_graph( g_len, periodic ? observe_every : 0, check_convergence, force_convergence ) )
            dev_step = true;
        else
        {
            (void)hipGetLastError();
            release_step_graph();
        }
    }
    const long whole = periodic ? ( n / observe_every ) * observe_every : 0; // steps in complete periods
    auto t0 = std::chrono::steady_clock::now();
    for( long it = 0; it < n; )
    {
        bool obs = ( it == n - 1 ) || ( observe_every > 0 && ( it + 1 ) % observe_every == 0 );
        if( dev_step )
        {
            bool replay = false;
            if( periodic )
                replay = it % observe_every == 0 && it + g_len <= whole;
            else if( !obs )
            {
                // steps it .. next_obs - 1 are unobserved
                long next_obs = n - 1;
                if( observe_every > 0 )
                    next_obs = std::min( next_obs, ( it / observe_every + 1 ) * (long)observe_every - 1 );
                replay = next_obs - it >= g_len;
            }
            if( replay )
            {
                HIPCHECK( hipGraphLaunch( step_graph, stream ) );
                it += g_len;
                step_counter += g_len;
                continue;
            }
        }
        step( obs, true, check_convergence && obs, force_convergence );
        ++it;
    }
    dev_step = false;
    eff_raw  = !min_is_solver( solver );
    HIPCHECK( hipGetLastError() );
    std::vector<long long> stop_step( n_img, 0 );
    HIPCHECK( hipMemcpyAsync( img_stopped.data(), b_img_stop.as<void>(), sizeof( int ) * n_img, hipMemcpyDeviceToHost,
                              stream ) );
    HIPCHECK( hipMemcpyAsync( stop_step.data(), b_stopstep.as<void>(), sizeof( long long ) * n_img,
                              hipMemcpyDeviceToHost, stream ) );
    HIPCHECK( hipMemcpyAsync( obs_host.data(), b_obs.as<void>(), sizeof( double ) * obs_host.size(),
                              hipMemcpyDeviceToHost, stream ) );
    HIPCHECK( hipStreamSynchronize( stream ) );
    auto t1      = std::chrono::steady_clock::now();
    last_step_ms = std::chrono::duration<double, std::milli>( t1 - t0 ).count() / (double)n;
    if( !pending.empty() )
        collect_profile();
    long done = 0;
    for( int k = 0; k < n_img; ++k )
    {
        if( was_stopped[k] )
            steps_last[k] = 0;
        else if( img_stopped[k] )
            steps_last[k] = (long)( stop_step[k] - (long long)first + 1 );
        else
            steps_last[k] = n;
        done = std::max( done, steps_last[k] );
    }
    step_counter = first + (uint64_t)done;
    return done;
}

void LLGRunner::tape_step( uint64_t step, double * s_in, double * s_out, double * F1, double * fv1, double * F2,
                           double * fv2 )
{
    if( solver != SOLVER_DEPONDT && solver != SOLVER_HEUN )
        throw std::runtime_error( "tape_step: Depondt or Heun only" );
    StepArgs A = STEP_ARGS();
    A.s        = s_in;
    A.s_out    = s_out;
    A.step     = step;
    A.obs      = 0;
    A.noise    = nullptr;
    HIPCHECK( hipMemsetAsync( b_stop.as<void>(), 0, sizeof( int ), stream ) );
    HIPCHECK( hipMemsetAsync( b_img_stop.as<void>(), 0, sizeof( int ) * specs.size(), stream ) );
    for( int st = 0; st < 2; ++st )
    {
        const double * x = st == 0 ? s_in : A.sp;
        if( any_ddi )
            ddi_field( x );
        A.tape_F  = st == 0 ? F1 : F2;
        A.tape_fv = st == 0 ? fv1 : fv2;
        llg_launch_stage( solver, st, A, n_blocks, stream );
    }
}

void LLGRunner::initial_observe()
{
    StepArgs A = STEP_ARGS();
    if( any_ddi )
        ddi_field( A.s );
    HIPCHECK( hipMemsetAsync( b_stop.as<void>(), 0, sizeof( int ), stream ) );
    llg_launch_observe( A, solver, n_blocks, stream );
    eff_raw = true;
    llg_launch_finalize( A, (int)specs.size(), b_obs.as<double>(), b_img_stop.as<int>(), b_nstop.as<int>(),
                         b_stopstep.as<long long>(), -1, 0.0, 0, stream );
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipMemcpyAsync( obs_host.data(), b_obs.as<void>(), sizeof( double ) * obs_host.size(),
                              hipMemcpyDeviceToHost, stream ) );
    HIPCHECK( hipStreamSynchronize( stream ) );
}

void LLGRunner::energy_terms( int img, double * out )
{
    StepArgs A = STEP_ARGS();
    if( any_ddi )
        ddi_field( A.s );
    DeviceBuffer e( sizeof( double ) * N_ENERGY_TERMS_HOST * (size_t)n_total );
    llg_launch_energy_terms( A, A.s, e.as<double>(), n_blocks, stream );
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipStreamSynchronize( stream ) );
    const ImageDev & I = imgs_host[img];
    const int nos      = specs[img].geometry->nos;
    HIPCHECK( hipMemcpy( out, e.as<double>() + N_ENERGY_TERMS_HOST * I.off, sizeof( double ) * N_ENERGY_TERMS_HOST * (size_t)nos,
                         hipMemcpyDeviceToHost ) );
}

double LLGRunner::gradient( int img, double * grad_aos )
{
    StepArgs A = STEP_ARGS();
    if( any_ddi )
        ddi_field( A.s );
    DeviceBuffer g( sizeof( double ) * 3 * (size_t)n_total );
    llg_launch_gradient( A, A.s, g.as<double>(), n_blocks, stream );
    HIPCHECK( hipGetLastError() );
    HIPCHECK( hipStreamSynchronize( stream ) );
    const ImageDev & I = imgs_host[img];
    const int nos      = specs[img].geometry->nos;
    std::vector<double> soa( 3 * (size_t)nos ), pE( I.n_blocks );
    HIPCHECK( hipMemcpy( soa.data(), g.as<double>() + 3 * I.off, sizeof( double ) * soa.size(),
                         hipMemcpyDeviceToHost ) );
    HIPCHECK( hipMemcpy( pE.data(), b_pE.as<double>() + I.block_begin, sizeof( double ) * I.n_blocks,
                         hipMemcpyDeviceToHost ) );
    if( grad_aos )
        soa_to_aos( soa.data(), grad_aos, nos );
    double E = 0;
    for( double v : pE )
        E += v;
    return E;
}

} // namespace spirit_hip
