// uts.cpp -- drop-in for `./uts <numThreads> <numElements> <fertility>
// <local_stack_size> <N_active_blocks>` (cuda/uts/uvm/uts.cu:550-587).
// Builds the same binary tree on the host (glibc srand(200), uts.cu:290-339),
// runs the GPU work-stealing search and verifies leaves, leaf sum and node
// count against the host golden (uts.cu:412-438); prints SUCCESS! or exits
// -1.  <numThreads> host threads take part, as the reference's pthreads
// were meant to (README.txt:30; its threads only meet at a barrier,
// uts.cu:457-539): this thread launches the GPU search and numThreads - 1
// host workers expand records the GPU hands them (hs_uts_ref_run_cpu).
// <local_stack_size> is accepted for compatibility (the GPU engine sizes its
// own per-workgroup deques); N_active_blocks is the number of workgroups
// (0 = default).
//
// Geometric SHA-1 trees:  ./uts --geo <seed> <b0> <depth> [numBlocks] [numThreads]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../../include/heterosync.h"

namespace {

// uts.cu:290-339: breadth-first "undecided" queue; a node is fertile with
// probability fertility% and gets two children; when the queue empties the
// current leaves become undecided again (most recent first).
void build_tree(int num_nodes, int fertility, std::vector<uint32_t> &value,
                std::vector<int32_t> &ca, std::vector<int32_t> &cb, uint32_t &leaves,
                uint32_t &leaf_sum, int &restarts) {
  value.assign(num_nodes, 0);
  ca.assign(num_nodes, 0);
  cb.assign(num_nodes, 0);
  std::vector<int32_t> undecided, leaf_list;  // leaf_list in creation order
  size_t head = 0;
  restarts = 0;
  srand(200);
  undecided.push_back(0);
  int next = 1;
  while (next != num_nodes) {
    if (head == undecided.size()) {
      undecided.assign(leaf_list.rbegin(), leaf_list.rend());
      head = 0;
      leaf_list.clear();
      ++restarts;
    }
    const int cur = undecided[head++];
    value[cur] = (uint32_t)(rand() % 100);
    if ((rand() % 100) < fertility) {
      ca[cur] = next;
      undecided.push_back(next++);
      cb[cur] = next;
      undecided.push_back(next++);
    } else {
      ca[cur] = cb[cur] = 0;
      leaf_list.push_back(cur);
    }
  }
  leaf_sum = 0;
  for (size_t i = head; i < undecided.size(); ++i) {
    ca[undecided[i]] = cb[undecided[i]] = 0;
    value[undecided[i]] = 1;
    leaf_sum += 1;
  }
  for (int32_t l : leaf_list) leaf_sum += value[l];
  leaves = (uint32_t)(leaf_list.size() + (undecided.size() - head));
}

}  // namespace

int main(int argc, char **argv) {
  hs_ctx *ctx = nullptr;
  if (argc >= 5 && !std::strcmp(argv[1], "--geo")) {
    const int seed = std::atoi(argv[2]);
    const double b0 = std::atof(argv[3]);
    const int depth = std::atoi(argv[4]);
    const int blocks = argc > 5 ? std::atoi(argv[5]) : 0;
    const int threads = argc > 6 ? std::atoi(argv[6]) : 1;
    if (threads < 1) { std::fprintf(stderr, "numThreads must be >= 1\n"); return -1; }
    if (hs_open(0, &ctx)) { std::fprintf(stderr, "no HIP device\n"); return -1; }
    hs_uts_result r;
    hs_uts_cpu_result cpu;
    const int rc = hs_uts_geo_run_cpu(ctx, seed, b0, depth, blocks, threads - 1, &r, &cpu);
    if (rc) { std::fprintf(stderr, "ERROR hs_uts_geo_run - %s\n", hs_strerror(rc)); return -1; }
    if (threads > 1)
      std::printf("host workers: %d, nodes expanded on the host: %llu (%llu batches)\n",
                  threads - 1, (unsigned long long)cpu.nodes,
                  (unsigned long long)cpu.deliveries);
    std::printf("Tree size = %llu, tree depth = %d, num leaves = %llu (%.2f%%)\n",
                (unsigned long long)r.nodes, r.max_depth, (unsigned long long)r.leaves,
                r.nodes ? 100.0 * r.leaves / r.nodes : 0.0);
    std::printf("kernel time: %f ms, %.3f Gnodes/s\n", r.kernel_ms,
                r.kernel_ms > 0 ? r.nodes / (r.kernel_ms * 1e6) : 0.0);
    hs_close(ctx);
    return 0;
  }
  if (argc != 6) {
    std::fprintf(stderr, "Invalid arguments passed in, expected:\n");
    std::fprintf(stderr, "./uts <numThreads> <numElements> <fertility> <local_stack_size> <N_active_blocks>\n");
    std::fprintf(stderr, "   or ./uts --geo <seed> <b0> <depth> [numBlocks] [numThreads]\n");
    return -1;
  }
  const int threads = std::atoi(argv[1]);
  const int N = std::atoi(argv[2]), fert = std::atoi(argv[3]), blocks = std::atoi(argv[5]);
  if (N <= 0) { std::fprintf(stderr, "N must be > 0!\n"); return -1; }
  if (threads < 1) { std::fprintf(stderr, "numThreads must be >= 1\n"); return -1; }
  const int num_nodes = N + ((N % 2) ? 0 : 1);  // odd (uts.cu:577)
  std::printf("Running with N_elements=%d N_active_blocks=%d\n", num_nodes, blocks);
  std::vector<uint32_t> value;
  std::vector<int32_t> ca, cb;
  uint32_t leaves = 0, leaf_sum = 0;
  int restarts = 0;
  build_tree(num_nodes, fert, value, ca, cb, leaves, leaf_sum, restarts);
  if (hs_open(0, &ctx)) { std::fprintf(stderr, "no HIP device\n"); return -1; }
  hs_uts_result r;
  hs_uts_cpu_result cpu;
  std::printf("Launch uts kernel...\n");
  const int rc = hs_uts_ref_run_cpu(ctx, num_nodes, value.data(), ca.data(), cb.data(),
                                    blocks > 0 ? blocks : 0, threads - 1, &r, &cpu);
  if (rc) { std::fprintf(stderr, "ERROR hs_uts_ref_run - %s\n", hs_strerror(rc)); return -1; }
  std::printf("Kernel completed.\n");
  if (threads > 1)
    std::printf("host workers: %d, nodes expanded on the host: %llu (%llu batches)\n",
                threads - 1, (unsigned long long)cpu.nodes, (unsigned long long)cpu.deliveries);
  std::printf("Created uts tree size %d with %u leaves, %u leaf sum, %d fertility, and %d restarts:\n",
              num_nodes, leaves, leaf_sum, fert, restarts);
  std::printf("kernel found %llu leaf nodes\n", (unsigned long long)r.leaves);
  std::printf("kernel found %llu leaf sum\n", (unsigned long long)r.leaf_sum);
  std::printf("nodes_processed:%llu\n", (unsigned long long)r.nodes);
  int bad = 0;
  if (r.leaves != leaves) { std::printf("Error: leaves_found %llu != %u\n", (unsigned long long)r.leaves, leaves); bad = 1; }
  if ((uint32_t)r.leaf_sum != leaf_sum) { std::printf("Error: leaf_sum %llu != %u\n", (unsigned long long)r.leaf_sum, leaf_sum); bad = 1; }
  if (r.nodes != (uint64_t)num_nodes) { std::printf("Error: nodes_processed %llu != %d\n", (unsigned long long)r.nodes, num_nodes); bad = 1; }
  hs_close(ctx);
  if (bad) return -1;
  std::printf("SUCCESS!\n");
  return 0;
}
