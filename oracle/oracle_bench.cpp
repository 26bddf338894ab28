// oracle_bench.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// Times the CPU restatement (zpic_oracle.hpp) on the em2d Weibel setup of the
// reference driver (openmp/em2d/main.cpp:179-259).  Used as a fallback
// cpu_baseline ("port") when the compiled reference driver is unavailable.
//
// usage: oracle_bench ntiles_x ntiles_y nx ny ppc_x ppc_y steps warmup

#include "zpic_oracle.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>

int main( int argc, char ** argv ) {
    int a[8] = { 8, 8, 16, 16, 8, 8, 5, 1 };
    for ( int i = 1; i < argc && i <= 8; i++ ) a[i - 1] = atoi( argv[i] );
    const int gx = a[0] * a[2], gy = a[1] * a[3];
    oracle::Sim<float> sim( a[0], a[1], a[2], a[3], (float) ( gx / 10.0 ), (float) ( gy / 10.0 ), (double) 0.07f );
    for ( int k = 0; k < 2; k++ ) {
        oracle::Species<float> s;
        s.m_q = k ? 1.0f : -1.0f;
        s.ppcx = a[4]; s.ppcy = a[5];
        s.ud.kind = 3;
        s.ud.uth[0] = s.ud.uth[1] = s.ud.uth[2] = 0.1f;
        s.ud.ufl[2] = k ? -0.6f : 0.6f;
        sim.add_species( s );
    }
    for ( int i = 0; i < a[7]; i++ ) sim.advance();
    uint64_t n0 = sim.sp[0].nmove + sim.sp[1].nmove;
    auto t0 = std::chrono::steady_clock::now();
    for ( int i = 0; i < a[6]; i++ ) sim.advance();
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>( t1 - t0 ).count();
    uint64_t moved = sim.sp[0].nmove + sim.sp[1].nmove - n0;
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    printf( "{\"particles_per_ns\": %.6g, \"ms_per_step\": %.6g, \"steps\": %d, \"threads\": %d}\n",
            moved / sec / 1e9, 1e3 * sec / a[6], a[6], threads );
    return 0;
}
