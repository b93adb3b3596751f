// stm_test.cpp — command line compatible with the reference's stm_test (src/main.cpp:11-63).
//
//   stm_test -v <method> -i <op file> [-i <op file> ...] [-t]
//     -v 38 | 39 | plain   root-to-leaf descent (the reference's STM versions)
//     -v learned           learned leaf routing
//     -i file              run the file's ops as batches ("R k", "P k v", "G a b"; signed
//                          keys and values, as the reference's Data_Collector<int, int>)
//     -t                   print every op's answer of the following -i files ("G a b": the
//                          number of pairs with a <= key <= b, then one "pair:" line each)
// Unlike the reference, -i before -v is an error instead of a crash (main.cpp:13,25-27),
// and -t switches printing on (the reference toggles a bool with ~, main.cpp:23).
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "gpu_btree.hpp"

int main(int argc, char** argv) {
  std::unique_ptr<stm_host::GPU_STM_Tree> tree;
  bool print = false;
  int opt;
  opterr = 0;
  try {
    while ((opt = getopt(argc, argv, "ti:v:")) != -1) {
      switch (opt) {
        case 't':
          print = true;
          break;
        case 'v': {
          uint32_t route = STM_ROUTE_DESCENT;
          if (std::strcmp(optarg, "learned") == 0) route = STM_ROUTE_LEARNED;
          else if (std::strcmp(optarg, "index") == 0) route = STM_ROUTE_INDEX;
          else if (std::strcmp(optarg, "38") && std::strcmp(optarg, "39") && std::strcmp(optarg, "plain")) {
            std::fprintf(stderr, "unknown method %s (38, 39, plain, learned or index)\n", optarg);
            return 2;
          }
          tree.reset(new stm_host::GPU_STM_Tree(route));
          tree->prepareGPU();
          break;
        }
        case 'i':
          if (!tree) {
            std::fprintf(stderr, "-i needs a tree: give -v first\n");
            return 2;
          }
          tree->prepareInput(optarg);
          tree->gpu_work();
          if (print) tree->test();
          break;
        default:
          std::fprintf(stderr, "usage: %s -v <38|39|plain|learned> -i <file> [-t]\n", argv[0]);
          return 2;
      }
    }
  } catch (const stm_host::StmFailure& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
