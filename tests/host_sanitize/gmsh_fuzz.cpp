// Host-only harness for the binary Gmsh 4.1 reader (femutils/Mesh.hpp),
// built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_host_sanitize.py.
//
//   gmsh_fuzz parse FILE...            parse each file (no device upload) and
//                                      print its node / cell / group counts
//   gmsh_fuzz mutate SEED ROUNDS FILE  parse ROUNDS corrupted copies of FILE
//                                      (truncations, byte flips, overwritten
//                                      64-bit counts): every copy must parse
//                                      or raise FatalError, never crash
//
// The parse-only path never calls the C ABI; the definitions below only
// satisfy the linker for the functions Mesh.hpp can reach on other paths.
#include "femutils/Mesh.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

extern "C" {
const char* afem_last_error(void) { return "host harness: no device library"; }
int afem_mesh_create(int, int32_t, const double*, int32_t, int32_t, const int32_t*, afem_mesh_t*)
{
  return AFEM_ERR_DEVICE;
}
int afem_mesh_destroy(afem_mesh_t) { return AFEM_OK; }
int afem_box_mesh_size(int, int, int32_t*, int32_t*) { return AFEM_ERR_DEVICE; }
int afem_box_mesh(int, int, double, uint64_t, double*, int32_t*) { return AFEM_ERR_DEVICE; }
}

using afem::femutils::FatalError;
using afem::femutils::Mesh;

static std::string read_file(const char* path)
{
  std::ifstream f(path, std::ios::binary);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

static uint64_t next(uint64_t& s) // splitmix64
{
  uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

int main(int argc, char** argv)
{
  if (argc >= 3 && std::string(argv[1]) == "parse") {
    for (int k = 2; k < argc; ++k) {
      Mesh m;
      m.readGmsh41(argv[k], false);
      std::printf("%s nodes=%d cells=%d npc=%d groups=%zu\n", argv[k], m.nbNode(), m.nbCell(), m.nodesPerCell(),
                  m.groups().size());
    }
    return 0;
  }
  if (argc == 5 && std::string(argv[1]) == "mutate") {
    uint64_t seed = std::strtoull(argv[2], nullptr, 10);
    const int rounds = std::atoi(argv[3]);
    const std::string orig = read_file(argv[4]);
    const std::string tmp = std::string(std::getenv("HARNESS_TMP") ? std::getenv("HARNESS_TMP") : "/tmp") +
                            "/gmsh_fuzz_" + std::to_string(seed) + ".msh";
    int parsed = 0, rejected = 0;
    for (int r = 0; r < rounds; ++r) {
      std::string d = orig;
      switch (next(seed) % 3) {
      case 0: // truncate
        d.resize(next(seed) % d.size());
        break;
      case 1: // flip bytes
        for (int q = 0; q < 8; ++q)
          d[next(seed) % d.size()] ^= (char)(1 + next(seed) % 255);
        break;
      default: { // overwrite an aligned 8-byte word with a huge or random count
        const size_t at = (next(seed) % (d.size() / 8)) * 8;
        const uint64_t v = (next(seed) & 1) ? ~0ULL - (next(seed) & 0xffff) : next(seed);
        std::memcpy(&d[at], &v, std::min<size_t>(8, d.size() - at));
      }
      }
      {
        std::ofstream o(tmp, std::ios::binary);
        o.write(d.data(), (std::streamsize)d.size());
      }
      try {
        Mesh m;
        m.readGmsh41(tmp, false);
        ++parsed;
      }
      catch (const FatalError&) {
        ++rejected;
      }
    }
    std::remove(tmp.c_str());
    std::printf("%s parsed=%d rejected=%d\n", argv[4], parsed, rejected);
    return 0;
  }
  std::fprintf(stderr, "usage: gmsh_fuzz parse FILE... | gmsh_fuzz mutate SEED ROUNDS FILE\n");
  return 2;
}
