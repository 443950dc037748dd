// `next` - NEXT-compatible driver over the GPU engine (include/nbx.h).
//
//   next <input.txt> <threads> <dt> <dump_interval> <vtk|vtu|hdf5> [options]
//
// The five positional arguments, the start-up lines, the loop and the
// snapshot cadence follow begrun.cpp:40-137 and argparse.cpp:21-61: each
// iteration takes dt from the current speeds (adaptive.h), makes one KDK
// step, advances the time, and writes dump_<k>.<ext> whenever the time has
// reached the next dump time.  Like NEXT, the loop has no end of its own; it
// stops on a 'q' line on stdin, or, with the options below, after a number of
// steps or at a time.
//
// Options (after the positional arguments):
//   --steps N        stop after N steps
//   --t-end T        stop once the simulated time reaches T
//   --theta X        opening angle (NEXT: 0.5, step.h:32)
//   --fixed-dt       use dt as given instead of the speed-adaptive step
//   --device D       HIP device ordinal
//   --quiet-dumps    run without writing snapshots (timing runs)
//   --dump-ic PATH   write the loaded particles to PATH in the chosen format
//                    and exit (no GPU use)
#include <poll.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "io.hpp"
#include "nbx.h"

namespace {

enum class Format { VTK, VTU, HDF5 };

struct Args {
  std::string input;
  int threads = 1;
  double dt = 0;
  double dump_interval = 0;
  Format format = Format::VTU;
  long long max_steps = -1;
  double t_end = -1;
  double theta = 0.5;
  bool fixed_dt = false;
  bool no_dumps = false;
  int device = 0;
  std::string dump_ic;
};

[[noreturn]] void usage_exit() {
  std::cerr << "Usage: next <input.txt> <threads> <dt> <dump_interval> <vtk|vtu|hdf5>\n";
  std::exit(1);
}

Args parse(int argc, char** argv) {
  if (argc < 6) usage_exit();
  Args a;
  try {
    a.input = argv[1];
    a.threads = std::stoi(argv[2]);
    a.dt = std::stod(argv[3]);
    a.dump_interval = std::stod(argv[4]);
  } catch (const std::exception&) {
    usage_exit();
  }
  const std::string fmt = argv[5];
  if (fmt == "vtk") a.format = Format::VTK;
  else if (fmt == "vtu") a.format = Format::VTU;
  else if (fmt == "hdf5") a.format = Format::HDF5;
  else {
    std::cerr << "Choose a file format: vtk, vtu, or hdf5\n";
    std::exit(1);
  }
  for (int i = 6; i < argc; ++i) {
    const std::string o = argv[i];
    auto value = [&]() -> std::string {
      if (i + 1 >= argc) { std::cerr << "missing value for " << o << "\n"; std::exit(1); }
      return argv[++i];
    };
    try {
      if (o == "--steps") a.max_steps = std::stoll(value());
      else if (o == "--t-end") a.t_end = std::stod(value());
      else if (o == "--theta") a.theta = std::stod(value());
      else if (o == "--device") a.device = std::stoi(value());
      else if (o == "--fixed-dt") a.fixed_dt = true;
      else if (o == "--quiet-dumps") a.no_dumps = true;
      else if (o == "--dump-ic") a.dump_ic = value();
      else { std::cerr << "unknown option " << o << "\n"; usage_exit(); }
    } catch (const std::exception&) {
      std::cerr << "bad value for " << o << "\n";
      std::exit(1);
    }
  }
  return a;
}

// A 'q' or 'Q' waiting on stdin ends the run (begrun.cpp:120-128).
bool quit_requested() {
  pollfd p{STDIN_FILENO, POLLIN, 0};
  if (poll(&p, 1, 0) <= 0 || !(p.revents & POLLIN)) return false;
  char buf[64];
  const ssize_t r = read(STDIN_FILENO, buf, sizeof buf);
  for (ssize_t k = 0; k < r; ++k)
    if (buf[k] == 'q' || buf[k] == 'Q') return true;
  return false;
}

bool save(const nextio::ParticleSet& ps, Format f, const std::string& path) {
  if (f == Format::VTK) return nextio::SaveVTK(ps, path);
  if (f == Format::VTU) return nextio::SaveVTU(ps, path);
#ifdef NBX_HDF5
  return nextio::SaveHDF5(ps, path);
#else
  return false;
#endif
}

int die(nbx_ctx* ctx, nbx_status s, const char* what) {
  std::cerr << "next: " << what << " failed (" << (int)s << "): " << nbx_last_error(ctx) << "\n";
  nbx_destroy(ctx);
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  Args args = parse(argc, argv);
#ifndef NBX_HDF5
  if (args.format == Format::HDF5) {
    std::cerr << "next: this build has no libhdf5; use vtk or vtu\n";
    return 1;
  }
#endif
  // begrun.cpp:60-67, trailing spaces included: a raw string that starts
  // and ends with a newline, printed with std::endl.
  static const char* kBanner =
      "\n"
      " _  _ ________   _________ \n"
      "| \\ | |  ____\\ \\ / /__   __|\n"
      "|  \\| | |__   \\ V /   | |   \n"
      "| . ` |  __|   > <    | |   \n"
      "| |\\  | |____ / . \\   | |   \n"
      "|_| \\_|______/_/ \\_\\  |_|   \n";
  std::cout << kBanner << std::endl;
  std::cout << " Threads:   " << args.threads << std::endl;
  std::cout << " Precision: FP64" << std::endl;

  nextio::ParticleSet ps = nextio::LoadParticles(args.input);
  std::cout << " Particles: " << ps.size() << std::endl;
  if (!args.dump_ic.empty()) {
    return save(ps, args.format, args.dump_ic) ? 0 : 1;
  }

  nbx_config cfg;
  nbx_config_default(&cfg);
  cfg.device = args.device;
  cfg.theta = args.theta;
  nbx_ctx* ctx = nullptr;
  nbx_status s = nbx_create(&ctx, &cfg);
  if (s != NBX_OK) {
    std::cerr << "next: no usable GPU (" << (int)s << ")\n";
    return 2;
  }
  const int64_t n = (int64_t)ps.size();
  s = nbx_upload(ctx, n, ps.x.data(), ps.y.data(), ps.z.data(), ps.vx.data(), ps.vy.data(), ps.vz.data(),
                 ps.m.data(), ps.type.data());
  if (s != NBX_OK) return die(ctx, s, "upload");

  double sim_time = 0, next_dump = 0;
  int dump = 0;
  long long steps = 0;
  const auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    if (args.max_steps >= 0 && steps >= args.max_steps) break;
    if (args.t_end >= 0 && sim_time >= args.t_end) break;
    double dt = args.dt;
    if (!args.fixed_dt) {
      s = nbx_adaptive_dt(ctx, args.dt, &dt);
      if (s != NBX_OK) return die(ctx, s, "adaptive dt");
    }
    s = nbx_step(ctx, dt);
    if (s != NBX_OK) return die(ctx, s, "step");
    sim_time += dt;
    ++steps;
    if (sim_time >= next_dump) {
      std::string out = "dump_" + std::to_string(dump);
      if (!args.no_dumps) {
        s = nbx_download(ctx, ps.x.data(), ps.y.data(), ps.z.data(), ps.vx.data(), ps.vy.data(), ps.vz.data());
        if (s != NBX_OK) return die(ctx, s, "download");
        out += args.format == Format::VTK ? ".vtk" : args.format == Format::VTU ? ".vtu" : ".hdf5";
        save(ps, args.format, out);
      }
      std::cout << "[Dump " << dump << "] t = " << sim_time << ", file: " << out << std::endl;
      next_dump += args.dump_interval;
      ++dump;
    }
    if (quit_requested()) {
      std::cout << "Exiting..." << std::endl;
      break;
    }
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cerr << "next: " << steps << " steps, t = " << sim_time << ", " << wall << " s" << std::endl;
  nbx_destroy(ctx);
  return 0;
}
