// prewrap.cpp -- see prewrap.h.
#include "prewrap.h"

#include <algorithm>
#include <vector>

namespace plprep {

plbn::Limbs raw(pleng::Group& G, const plbn::Modulus& M, uint32_t slot)
{
    std::vector<int32_t> d;
    int64_t xhi = 0;
    G.download(slot, d, xhi);
    return plbn::from_digits(M, G.plan().L, d, xhi);
}

void set_raw(pleng::Group& G, const plbn::Modulus& M, uint32_t slot, const plbn::Limbs& value)
{
    std::vector<int32_t> d;
    int64_t xhi = 0;
    plbn::to_digits(M, G.plan().L, plbn::mod_N(M, value), d, xhi);
    G.upload(slot, d, xhi);
}

void load_constants(pleng::Group& G, const plbn::Modulus& M, uint64_t a, const Scratch& s)
{
    const plbn::Limbs c = plbn::kinv_pow(M, 4);
    set_raw(G, M, s.m1, c);
    if (a) set_raw(G, M, s.ma, plbn::mul_small_mod(M, c, a));
}

double step(pleng::Group& G, uint32_t z, const Scratch& s, bool times_a)
{
    double e = G.square_slots(z, 1, 1, nullptr);  // z^2 * k^2
    G.copy_slot(s.t, times_a ? s.ma : s.m1);       // the product leaves its operand transformed
    return std::max(e, G.mul_slot(z, s.t));        // * m * k^-4 * k^2
}

double to_mont(pleng::Group& G, uint32_t z, const Scratch& s)
{
    G.copy_slot(s.t, s.m1);
    return G.mul_slot(z, s.t);  // z * k^-4 * k^2 = z * 2^(2n)
}

uint64_t wrap_bits(const plbn::Modulus& M) { return (uint64_t)M.n + plbn::bitlen_u64(M.k) + 16; }

uint64_t square_bits(uint64_t bits, uint64_t a)
{
    const uint64_t cap = 1ull << 62;
    if (bits >= cap / 2) return cap;
    return 2 * bits + (a ? plbn::bitlen_u64(a) : 0);
}

bool SmallValue::set(const plbn::Modulus& M, const plbn::Limbs& x0)
{
    const plbn::Limbs x = plbn::mod_N(M, x0);
    bits = plbn::distance_bits(M, x);
    val = 0;
    if (bits <= 32) {
        const bool low =
            std::all_of(x.begin() + 1, x.end(), [](uint64_t v) { return v == 0; }) && plbn::bitlen_u64(x[0]) <= 32;
        plbn::Limbs t = M.N;
        plbn::sub_in(t, x);
        val = low ? x[0] : t[0];
    }
    return bits + 64 < M.n;
}

void SmallValue::grow(bool square, uint64_t m)
{
    if (bits <= 32) {
        unsigned __int128 v = val;
        if (square) v *= v;
        if (m) v *= m;
        if (v < ((unsigned __int128)1 << 62)) {
            val = (uint64_t)v;
            bits = plbn::bitlen_u64(val);
            return;
        }
    }
    bits = square ? square_bits(bits, m) : bits + plbn::bitlen_u64(m);
    val = 0;
}

}  // namespace plprep
