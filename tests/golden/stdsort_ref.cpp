// Reads n uint64 keys (binary, stdin), sorts them with libstdc++ std::sort
// comparing only the high 32 bits (so equal high halves expose the
// algorithm's tie order) and writes them back (binary, stdout).  Used by the
// CPU tests to pin the oracle's restatement of std::sort.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

int main() {
    std::vector<uint64_t> v;
    uint64_t x;
    while (std::fread(&x, sizeof x, 1, stdin) == 1) v.push_back(x);
    std::sort(v.begin(), v.end(), [](uint64_t a, uint64_t b) { return (a >> 32) < (b >> 32); });
    std::fwrite(v.data(), sizeof(uint64_t), v.size(), stdout);
    return 0;
}
