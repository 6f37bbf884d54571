// Feeds length-prefixed inputs from a file to the plan JSON parser
// (proteus_amd/engine/json.cpp) and round-trips the accepted ones through
// dump(); built with AddressSanitizer and UBSan by tests/test_json_fuzz.py.
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../proteus_amd/engine/json.hpp"

int main(int argc, char **argv) {
  if (argc != 2) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  unsigned long accepted = 0, rejected = 0;
  for (;;) {
    uint32_t len = 0;
    if (std::fread(&len, 4, 1, f) != 1) break;
    std::string text(len, '\0');
    if (len && std::fread(&text[0], 1, len, f) != len) return 3;
    pj::engine::Json j;
    std::string err;
    if (pj::engine::json_parse(text, &j, &err)) {
      ++accepted;
      pj::engine::Json k;
      const std::string d = j.dump();
      if (!pj::engine::json_parse(d, &k, &err)) {  // dump must re-parse
        std::fprintf(stderr, "dump does not re-parse (%s)\ninput: %.200s\ndump: %.200s\n",
                     err.c_str(), text.c_str(), d.c_str());
        return 4;
      }
    } else {
      ++rejected;
    }
  }
  std::fclose(f);
  std::printf("accepted %lu rejected %lu\n", accepted, rejected);
  return 0;
}
