// bignum.h -- exact host-side arithmetic modulo N = k*2^n + 1 on
// little-endian u64 limbs.  Used only at the edges of a test (seed,
// Montgomery conversions, canonical residue, Res64); the squaring loop never
// leaves the device.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace plbn {

using Limbs = std::vector<uint64_t>;

struct Modulus {
    uint32_t n;
    uint64_t k;
    size_t nl;  // limbs that hold any value < 2^(n+128)
    Limbs N;    // k*2^n+1 over nl limbs
    Modulus(uint32_t n_, uint64_t k_);
};

int cmp(const Limbs& a, const Limbs& b);            // same length
void add_in(Limbs& a, const Limbs& b);                // a += b (same length, carry dropped)
void sub_in(Limbs& a, const Limbs& b);                // a -= b (same length, borrow dropped)
uint64_t divmod_small(Limbs& a, uint64_t d);          // a /= d, returns a mod d
void mul_small_in(Limbs& a, uint64_t m);              // a *= m (a must have head-room)
void shift_left_bits(const Limbs& a, unsigned s, Limbs& out);

// x (any length, any size) -> x mod N over M.nl limbs.
Limbs mod_N(const Modulus& M, Limbs x);
// x*2^n mod N for x < N.
Limbs mul_2n(const Modulus& M, const Limbs& x);
// Montgomery forms for R = 2^(2n):  to: x*2^(2n),  from: y*k^2  (2^-n == -k).
Limbs to_mont(const Modulus& M, const Limbs& x);
Limbs from_mont(const Modulus& M, const Limbs& y);
Limbs mul_small_mod(const Modulus& M, const Limbs& x, uint64_t a);

// Variable-base balanced digits: digit m sits at bit e_m = ceil(n*m/L).
// to_digits: x in [0, N) -> digits[0..L) in [-2^(b-1), 2^(b-1)) and the
// top scalar xhi = x >> n plus the last carry.
void to_digits(const Modulus& M, uint32_t L, const Limbs& x, std::vector<int32_t>& digits, int64_t& xhi);
// from_digits: sum digits[m]*2^e_m + xhi*2^n, reduced to [0, N).
Limbs from_digits(const Modulus& M, uint32_t L, const std::vector<int32_t>& digits, int64_t xhi);

// x*k^-1 mod N for x < N (k^-1 == (k-1)*2^n + 1, so one shift-fold and a
// small multiply), and k^-e mod N.
Limbs mul_kinv(const Modulus& M, const Limbs& x);
Limbs kinv_pow(const Modulus& M, unsigned e);
// Bit length of min(x, N - x) for x < N: how far a residue is from the
// small values whose digits are sparse.
uint64_t distance_bits(const Modulus& M, const Limbs& x);

uint64_t low64(const Limbs& x);
bool is_n_minus_1(const Modulus& M, const Limbs& x);
unsigned bitlen_u64(uint64_t x);

}  // namespace plbn
