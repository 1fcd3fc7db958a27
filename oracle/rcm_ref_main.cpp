// Runs the reference's own genrcm (src/math/RCM.cpp:1214-1327, compiled in
// place by oracle/ref.mk) on a graph given as 0-based CSR in a binary file:
//   int32 n, int32 nnz, int32 ptr[n+1], int32 adj[nnz]
// and writes the 0-based permutation perm[new] = old as n int32.
//
// Test infrastructure only: pins the oracle's and the library's GENRCM on
// arbitrary graphs (tests/test_lib_cpu.py).
#include <cstdio>
#include <vector>

void genrcm(int node_num, int adj_num, int adj_row[], int adj[], int perm[]);

int main(int argc, char** argv) {
    if (argc != 3) { fprintf(stderr, "usage: %s graph.bin perm.bin\n", argv[0]); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0, m = 0;
    if (fread(&n, 4, 1, f) != 1 || fread(&m, 4, 1, f) != 1) return 2;
    std::vector<int> ptr(n + 1), adj(m > 0 ? m : 1), perm(n);
    if (fread(ptr.data(), 4, n + 1, f) != (size_t)(n + 1) || fread(adj.data(), 4, m, f) != (size_t)m) return 2;
    fclose(f);
    for (auto& p : ptr) p += 1;          // the reference works 1-based
    for (int k = 0; k < m; ++k) adj[k] += 1;
    genrcm(n, m, ptr.data(), adj.data(), perm.data());
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 2;
    for (auto& p : perm) p -= 1;
    fwrite(perm.data(), 4, n, g);
    fclose(g);
    return 0;
}
