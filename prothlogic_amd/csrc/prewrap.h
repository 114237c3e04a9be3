// prewrap.h -- the first squarings of a test, while its exact value is small.
//
// The device holds a residue x as the state y = x*2^(2n) mod N (the two
// folds of a squaring divide by 2^(2n)).  2^(2n) == k^-2 (mod N), and for k
// whose powers of two repeat soon modulo k^2 (k = 3, 5, 7, 15, 17, 2^m - 1,
// 2^m + 1, ...) the binary expansion of k^-2 mod N is periodic.  Small x then
// give states x*k^-2 mod N whose balanced digits are nearly all alike, and
// the squared transform of such coherent digits has coefficients about
// sqrt(L) times larger than for random digits: 3*2^303093+1 at L = 2^14
// reaches round-off 0.44 in the first 15 squarings and 0.02 afterwards.
//
// So while the exact value a^(k*2^i) (or the seed a^j) has fewer bits than
// N, the slot holds x itself ("normal form"): its digits are sparse, or
// random once the value has wrapped modulo N.  A squaring in normal form is
// two device products, D(D(z, z), c) with D(u, v) = u*v*2^(-2n) = u*v*k^2
// and c = m*k^-4 (m = a for a seed step whose bit of k is set, else 1):
// D(z, z) = z^2*k^2, and D(z^2*k^2, m*k^-4) = m*z^2.  The constant c is
// structured but each product pairs it with a sparse or random operand,
// which keeps the coefficients at the random-digit size.  Once the exact
// value passes wrap_bits() one more product D(z, k^-4) = z*k^-2 = z*2^(2n)
// moves the slot to the Montgomery form the steady squarings use.
#pragma once
#include <cstdint>

#include "bignum.h"
#include "engine.h"

namespace plprep {

// Slots a normal-form run uses beside the residue's own: a temporary and
// the two constants.
struct Scratch {
    uint32_t t = 0, m1 = 0, ma = 0;
};

// Canonical value of a slot's raw state, and a raw state from a value.
plbn::Limbs raw(pleng::Group& G, const plbn::Modulus& M, uint32_t slot);
void set_raw(pleng::Group& G, const plbn::Modulus& M, uint32_t slot, const plbn::Limbs& value);

// Loads k^-4 into s.m1 and a*k^-4 into s.ma (a = 0: s.m1 only).
void load_constants(pleng::Group& G, const plbn::Modulus& M, uint64_t a, const Scratch& s);
// z <- z^2 * (times_a ? a : 1), z in normal form; returns the round-off.
double step(pleng::Group& G, uint32_t z, const Scratch& s, bool times_a);
// Normal form -> Montgomery form of slot z; returns the round-off.
double to_mont(pleng::Group& G, uint32_t z, const Scratch& s);
// Bit length of the exact value beyond which its residue mod N is taken as
// random (2^16 * N: the low n bits of the value decide the residue).
uint64_t wrap_bits(const plbn::Modulus& M);
// Bit-length bound of x^2 (or x^2 * a) from that of x, saturating.
uint64_t square_bits(uint64_t bits, uint64_t a);

// Size of a residue x while it is small: a bound on the bit length of
// min(x, N - x), and that value itself while the bound is at most 32 bits
// (so 0 and 1, which squaring leaves alone, are known to stay small).
struct SmallValue {
    uint64_t bits = 0;
    uint64_t val = 0;
    // true when x is small enough for the normal form
    bool set(const plbn::Modulus& M, const plbn::Limbs& x);
    // x <- x*m, or x <- x^2*m with square (m = 0 meaning 1)
    void grow(bool square, uint64_t m);
    bool wrapped(const plbn::Modulus& M) const { return bits > wrap_bits(M); }
};

}  // namespace plprep
