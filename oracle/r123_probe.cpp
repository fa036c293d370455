// r123_probe.cpp -- test infrastructure only.
// Evaluates the reference's own Random123 code (Philox4x32_R<7|10> from
// Random123/include/Random123/philox.h and u01fixedpt<float> from
// Random123/examples/uniform.hpp) on queries read from stdin, so that the
// oracle's restatement can be compared with the reference implementation.
// Built by oracle/Makefile into oracle/_ref/ with the reference's include
// directories; it is never part of the product.
//
// stdin lines:
//   P R c0 c1 c2 c3 k0 k1   -> 4 hex words of Philox4x32_R<R>(ctr, key)
//   U r                     -> hex bits of u01fixedpt<float>(r)
#include <Random123/philox.h>
#include <uniform.hpp>
#include <cstdio>
#include <cstring>
#include <cstdint>

template <unsigned R> static void run(const uint32_t c[4], const uint32_t k[2]) {
    typedef r123::Philox4x32_R<R> G;
    G g;
    typename G::ctr_type ctr = {{c[0], c[1], c[2], c[3]}};
    typename G::key_type key = {{k[0], k[1]}};
    typename G::ctr_type r = g(ctr, key);
    std::printf("%08x %08x %08x %08x\n", r.v[0], r.v[1], r.v[2], r.v[3]);
}

int main() {
    char tag[4];
    while (std::scanf("%3s", tag) == 1) {
        if (tag[0] == 'P') {
            unsigned R;
            uint32_t c[4], k[2];
            if (std::scanf("%u %x %x %x %x %x %x", &R, &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) != 7) return 2;
            if (R == 7) run<7>(c, k);
            else if (R == 10) run<10>(c, k);
            else return 3;
        } else if (tag[0] == 'U') {
            uint32_t r;
            if (std::scanf("%x", &r) != 1) return 2;
            float f = r123::u01fixedpt<float>(r);
            uint32_t b;
            std::memcpy(&b, &f, 4);
            std::printf("%08x\n", b);
        } else {
            return 4;
        }
    }
    return 0;
}
