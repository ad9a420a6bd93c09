// main.cpp -- the cabanamd command-line driver (bin/main.cpp:59-78 with the
// options of InputCL::read_args, inputCL.cpp:69-258), running the HIP engine
// through the C++ classes of cabanamd_amd.hpp.
//
//   cabanamd -il in.lj [--force-iteration NEIGH_FULL|NEIGH_HALF]
//            [-o out.txt] [-e err.txt] [--precision FP64|MIXED] [--lanes N]
//
// Output follows CbnMD::print_summary and the final timing block of
// CbnMD::run (cabanamd_impl.h:400-432,644-676).
#include "cabanamd_amd.hpp"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <array>
#include <string>

namespace
{
struct Options
{
    std::string input, output, error;
    int force_iteration = -1; // deck decides (newton on/off)
    int precision = CMD_PREC_FP64;
    int lanes = 0; // 0: autotune
    int dump_rate = 0, check_rate = 0; // --dumpbinary / --correctness
    double vacuum_rate = 1.0;          // --vacuum
    std::string dump_path, ref_path, check_file;
};

void usage( std::ostream &o )
{
    o << "cabanamd (MI355X HIP engine)\n"
         "  -il, --input-lammps FILE   LAMMPS-subset input deck (required)\n"
         "  -o,  --output-file FILE    append output to FILE as well\n"
         "  -e,  --error-file FILE     append errors to FILE as well\n"
         "  --force-iteration NEIGH_FULL|NEIGH_HALF\n"
         "  --neigh-type VERLET_2D|VERLET_CSR\n"
         "  --neigh-parallel SERIAL|TEAM|TEAM_VECTOR (accepted; wavefront mapping)\n"
         "  --device-type HIP\n"
         "  --precision FP64|MIXED     pair arithmetic (default FP64)\n"
         "  --lanes 0|1|2|4|8          lanes per atom, 0 = autotune\n"
         "  --dumpbinary N PATH        snapshot every N steps to PATH/output.*\n"
         "  --correctness N PATH FILE  compare with PATH/output.* every N steps,\n"
         "                             differences appended to FILE\n"
         "  --vacuum N                 scale the box's high corner by N > 1\n";
}

[[noreturn]] void fail( const std::string &msg )
{
    throw std::runtime_error( msg );
}

Options parse( int argc, char **argv )
{
    Options o;
    auto value = [&]( int &i ) -> std::string
    {
        if ( i + 1 >= argc )
            fail( std::string( "Missing value for commandline option: " ) + argv[i] );
        return argv[++i];
    };
    for ( int i = 1; i < argc; ++i )
    {
        const std::string a = argv[i];
        if ( a == "-h" || a == "--help" )
        {
            usage( std::cout );
            std::exit( 0 );
        }
        else if ( a == "-il" || a == "--input-lammps" )
            o.input = value( i );
        else if ( a == "-o" || a == "--output-file" )
            o.output = value( i );
        else if ( a == "-e" || a == "--error-file" )
            o.error = value( i );
        else if ( a == "--device-type" )
        {
            const std::string v = value( i );
            if ( v != "HIP" )
                fail( "Unknown commandline option: " + a + " " + v +
                      " (this build runs on HIP devices only)" );
        }
        else if ( a == "--force-iteration" )
        {
            const std::string v = value( i );
            if ( v == "NEIGH_FULL" )
                o.force_iteration = 0;
            else if ( v == "NEIGH_HALF" )
                o.force_iteration = 1;
            else
                fail( "Unknown commandline option: " + a + " " + v );
        }
        else if ( a == "--neigh-type" )
        {
            const std::string v = value( i );
            if ( v != "VERLET_2D" && v != "VERLET_CSR" )
                fail( "Unknown commandline option: " + a + " " + v );
        }
        else if ( a == "--neigh-parallel" )
        {
            const std::string v = value( i );
            if ( v != "SERIAL" && v != "TEAM" && v != "TEAM_VECTOR" )
                fail( "Unknown commandline option: " + a + " " + v );
        }
        else if ( a == "--precision" )
        {
            const std::string v = value( i );
            if ( v == "FP64" )
                o.precision = CMD_PREC_FP64;
            else if ( v == "MIXED" )
                o.precision = CMD_PREC_MIXED;
            else
                fail( "Unknown commandline option: " + a + " " + v );
        }
        else if ( a == "--dumpbinary" )
        {
            o.dump_rate = std::stoi( value( i ) );
            o.dump_path = value( i );
            if ( o.dump_rate <= 0 )
                fail( "--dumpbinary needs a positive rate" );
        }
        else if ( a == "--correctness" )
        {
            o.check_rate = std::stoi( value( i ) );
            o.ref_path = value( i );
            o.check_file = value( i );
            if ( o.check_rate <= 0 )
                fail( "--correctness needs a positive rate" );
        }
        else if ( a == "--lanes" )
            o.lanes = std::stoi( value( i ) );
        else if ( a == "--vacuum" )
        {
            // inputCL.cpp:243-250
            o.vacuum_rate = std::atof( value( i ).c_str() );
            if ( o.vacuum_rate <= 1.0 )
                fail( "Vacuum multiplier must be bigger than 1.0" );
        }
        else
            fail( "Unknown commandline option: " + a );
    }
    if ( o.input.empty() )
        fail( "No input file given (-il FILE)" );
    return o;
}

// Writes to stdout and, when given, appends to the output file.
struct Log
{
    std::ofstream file;
    template <class T>
    Log &operator<<( const T &v )
    {
        std::cout << v;
        if ( file.is_open() )
            file << v;
        return *this;
    }
};

std::string fmt( const char *f, double v )
{
    char b[64];
    std::snprintf( b, sizeof( b ), f, v );
    return b;
}
} // namespace

int main( int argc, char **argv )
{
    Options o;
    try
    {
        o = parse( argc, argv );
    }
    catch ( const std::exception &e )
    {
        std::cerr << e.what() << "\n";
        usage( std::cerr );
        return 1;
    }
    Log out;
    if ( !o.output.empty() )
        out.file.open( o.output, std::ofstream::app );
    try
    {
        using namespace cabanamd_amd;
        check( cmd_device_init( 0 ) );
        InputFile input;
        input.read_file( o.input.c_str(), o.force_iteration );
        input.params.vacuum_rate = o.vacuum_rate;
        const cmd_params &p = input.params;
        CbnMD md( p );
        md.set_precision( o.precision );
        md.set_lanes( o.lanes );
        md.init();
        // phase times (T_Force, T_Neigh, T_Comm, T_Int) from HIP events
        check( cmd_md_profile( md.handle(), 1 ) );
        const int thermo = p.thermo_rate > 0 ? p.thermo_rate : p.nsteps;
        auto rows = md.thermo();
        int N = 0;
        {
            cmd_system *s = nullptr;
            check( cmd_md_system( md.handle(), &s ) );
            int nl, ng;
            check( cmd_system_counts( s, &N, &nl, &ng ) );
        }
        // print_summary (cabanamd_impl.h:647-668): step, then T, PE/N and
        // E/N fixed with 6 digits, time fixed with 2, the rate scientific
        // with 2
        auto row = [&]( int step, const std::array<double, 5> &r, double time, double rate )
        {
            out << step << "\t" << fmt( "%.6f", r[1] ) << "\t" << fmt( "%.6f", r[2] ) << "\t"
                << fmt( "%.6f", r[3] ) << "\t" << fmt( "%.2f", time ) << "\t"
                << fmt( "%.2e", rate ) << "\n";
        };
        out << "\n#Timestep Temperature PotE ETot Time Atomsteps/s \n";
        if ( !rows.empty() )
            row( 0, rows[0], 0.0, 0.0 );
        // dump_binary / check_correctness at step 0 and every rate steps
        // (cabanamd_impl.h:245-251, 392-396)
        auto at_step = [&]( int step )
        {
            if ( o.dump_rate > 0 && step % o.dump_rate == 0 )
                check( cmd_md_dump( md.handle(), o.dump_path.c_str(), step ) );
            if ( o.check_rate > 0 && step % o.check_rate == 0 )
            {
                double d[6];
                check( cmd_md_check_correctness( md.handle(), o.ref_path.c_str(), step,
                                                 o.check_file.c_str(), d ) );
            }
        };
        at_step( 0 );
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        double last = 0.0, last_row = 0.0;
        int done = 0;
        while ( done < p.nsteps )
        {
            int chunk = std::min( thermo - done % thermo, p.nsteps - done );
            if ( o.dump_rate > 0 )
                chunk = std::min( chunk, o.dump_rate - done % o.dump_rate );
            if ( o.check_rate > 0 )
                chunk = std::min( chunk, o.check_rate - done % o.check_rate );
            md.run( chunk );
            done += chunk;
            at_step( done );
            const double now =
                std::chrono::duration<double>( clock::now() - t0 ).count();
            rows = md.thermo();
            if ( done % thermo == 0 && !rows.empty() )
            {
                // rate over the steps since the last row (cabanamd_impl.h:367-368)
                row( done, rows.back(), now, 1.0 * N * thermo / ( now - last_row ) );
                last_row = now;
            }
            last = now;
        }
        // the final block (cabanamd_impl.h:400-421): one process, no load
        // balancer (T_lb 0), everything not in a phase is T_Other
        const double time = last;
        const auto t = md.timers();
        const double other = std::max( 0.0, time - t[0] - t[1] - t[2] - t[3] );
        const double ph[6] = { t[0], t[1], t[2], t[3], 0.0, other };
        out << "\n#Procs Atoms | Time T_Force T_Neigh T_Comm T_Int T_lb T_Other |\n"
            << 1 << " " << N << " | " << fmt( "%.2f", time );
        for ( double v : ph )
            out << " " << fmt( "%.2f", v );
        out << " | PERFORMANCE\n" << 1 << " " << N << " | " << fmt( "%.2f", 1.0 );
        for ( double v : ph )
            out << " " << fmt( "%.2f", time > 0 ? v / time : 0.0 );
        out << " | FRACTION\n\n";
        const double sps = time > 0 ? p.nsteps / time : 0.0;
        out << "#Steps/s Atomsteps/s Atomsteps/(proc*s)\n" << fmt( "%.2e", sps ) << " "
            << fmt( "%.2e", sps * N ) << " " << fmt( "%.2e", sps * N ) << "\n";
    }
    catch ( const std::exception &e )
    {
        std::cerr << e.what() << "\n";
        if ( !o.error.empty() )
        {
            std::ofstream err( o.error, std::ofstream::app );
            err << e.what() << "\n";
        }
        return 1;
    }
    return 0;
}
