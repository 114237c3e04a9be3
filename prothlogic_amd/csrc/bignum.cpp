// bignum.cpp -- see bignum.h.
#include "bignum.h"

#include <algorithm>
#include <stdexcept>

namespace plbn {

using u128 = unsigned __int128;

unsigned bitlen_u64(uint64_t x)
{
    unsigned b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b;
}

Modulus::Modulus(uint32_t n_, uint64_t k_) : n(n_), k(k_)
{
    if (k_ == 0 || (k_ & 1) == 0) throw std::invalid_argument("k must be odd and positive");
    nl = (size_t)(n_ + 128) / 64 + 2;
    N.assign(nl, 0);
    N[0] = 1;
    size_t q = n_ / 64;
    unsigned r = n_ % 64;
    N[q] += (r == 0) ? k_ : (k_ << r);
    if (r) N[q + 1] += k_ >> (64 - r);
}

int cmp(const Limbs& a, const Limbs& b)
{
    size_t n = std::max(a.size(), b.size());
    for (size_t i = n; i-- > 0;) {
        uint64_t x = i < a.size() ? a[i] : 0, y = i < b.size() ? b[i] : 0;
        if (x != y) return x > y ? 1 : -1;
    }
    return 0;
}

void add_in(Limbs& a, const Limbs& b)
{
    uint64_t c = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        u128 s = (u128)a[i] + (i < b.size() ? b[i] : 0) + c;
        a[i] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
    }
}

void sub_in(Limbs& a, const Limbs& b)
{
    uint64_t br = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        uint64_t y = i < b.size() ? b[i] : 0;
        u128 d = (u128)a[i] - y - br;
        a[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) ? 1 : 0;
    }
}

uint64_t divmod_small(Limbs& a, uint64_t d)
{
    uint64_t rem = 0;
    for (size_t i = a.size(); i-- > 0;) {
        u128 num = ((u128)rem << 64) | a[i];
        a[i] = (uint64_t)(num / d);
        rem = (uint64_t)(num % d);
    }
    return rem;
}

void mul_small_in(Limbs& a, uint64_t m)
{
    uint64_t c = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        u128 p = (u128)a[i] * m + c;
        a[i] = (uint64_t)p;
        c = (uint64_t)(p >> 64);
    }
    if (c) a.push_back(c);
}

static void trim(Limbs& a)
{
    while (a.size() > 1 && a.back() == 0) a.pop_back();
}

static Limbs pow2n_times(const Modulus& M, uint64_t r)  // r*2^n over nl limbs
{
    Limbs A(M.nl, 0);
    size_t q = M.n / 64;
    unsigned s = M.n % 64;
    A[q] = (s == 0) ? r : (r << s);
    if (s) A[q + 1] = r >> (64 - s);
    return A;
}

Limbs mod_N(const Modulus& M, Limbs x)
{
    trim(x);
    if (x.size() < M.nl) x.resize(M.nl, 0);
    if (cmp(x, M.N) < 0) {
        x.resize(M.nl);
        return x;
    }
    // x = H*2^n + Lo, H = q*k + r  =>  x == Lo + r*2^n - q  (k*2^n == -1)
    size_t q = M.n / 64;
    unsigned s = M.n % 64;
    Limbs H(x.size() - q + 1, 0);
    for (size_t i = 0; i + q < x.size(); ++i) {
        uint64_t h = x[i + q] >> s;
        if (s && i + q + 1 < x.size()) h |= x[i + q + 1] << (64 - s);
        H[i] = h;
    }
    uint64_t r = divmod_small(H, M.k);
    Limbs A(M.nl, 0);
    for (size_t i = 0; i < q; ++i) A[i] = x[i];
    if (s) A[q] = x[q] & ((1ULL << s) - 1);
    add_in(A, pow2n_times(M, r));  // A < N
    Limbs Q = mod_N(M, H);
    if (cmp(A, Q) >= 0) {
        sub_in(A, Q);
    } else {
        add_in(A, M.N);
        sub_in(A, Q);
    }
    while (cmp(A, M.N) >= 0) sub_in(A, M.N);
    return A;
}

Limbs mul_2n(const Modulus& M, const Limbs& x)
{
    Limbs q = x;
    q.resize(M.nl, 0);
    uint64_t r = divmod_small(q, M.k);  // x = q*k + r ; x*2^n == r*2^n - q
    Limbs A = pow2n_times(M, r);
    if (cmp(A, q) >= 0) {
        sub_in(A, q);
    } else {
        add_in(A, M.N);
        sub_in(A, q);
    }
    return A;
}

Limbs to_mont(const Modulus& M, const Limbs& x)
{
    Limbs y = mod_N(M, x);
    return mul_2n(M, mul_2n(M, y));
}

Limbs from_mont(const Modulus& M, const Limbs& y)
{
    Limbs v = y;
    v.resize(M.nl + 2, 0);
    mul_small_in(v, M.k);
    mul_small_in(v, M.k);
    return mod_N(M, v);
}

Limbs mul_small_mod(const Modulus& M, const Limbs& x, uint64_t a)
{
    Limbs v = x;
    v.resize(M.nl + 1, 0);
    mul_small_in(v, a);
    return mod_N(M, v);
}

static inline uint64_t digit_pos(uint64_t n, uint64_t L, uint64_t m)  // ceil(n*m/L)
{
    return (n * m + L - 1) / L;
}

static inline uint64_t get_bits(const Limbs& x, uint64_t pos, unsigned w)
{
    size_t li = pos / 64;
    unsigned off = pos % 64;
    uint64_t v = li < x.size() ? x[li] >> off : 0;
    if (off && off + w > 64 && li + 1 < x.size()) v |= x[li + 1] << (64 - off);
    return w >= 64 ? v : (v & ((1ULL << w) - 1));
}

static inline void or_bits(Limbs& x, uint64_t pos, unsigned w, uint64_t v)
{
    size_t li = pos / 64;
    unsigned off = pos % 64;
    x[li] |= v << off;
    if (off && off + w > 64) x[li + 1] |= v >> (64 - off);
}

void to_digits(const Modulus& M, uint32_t L, const Limbs& x, std::vector<int32_t>& digits, int64_t& xhi)
{
    digits.assign(L, 0);
    uint64_t carry = 0;
    for (uint32_t m = 0; m < L; ++m) {
        uint64_t e = digit_pos(M.n, L, m), e1 = digit_pos(M.n, L, m + 1);
        unsigned b = (unsigned)(e1 - e);
        uint64_t u = get_bits(x, e, b) + carry;
        uint64_t half = 1ULL << (b - 1);
        if (u >= half) {
            digits[m] = (int32_t)((int64_t)u - (int64_t)(1ULL << b));
            carry = 1;
        } else {
            digits[m] = (int32_t)u;
            carry = 0;
        }
    }
    // x >> n: below 2^64 for x < N
    uint64_t top = get_bits(x, M.n, 64);
    xhi = (int64_t)top + (int64_t)carry;
}

Limbs from_digits(const Modulus& M, uint32_t L, const std::vector<int32_t>& digits, int64_t xhi)
{
    Limbs out(M.nl, 0);
    int64_t carry = 0;
    for (uint32_t m = 0; m < L; ++m) {
        uint64_t e = digit_pos(M.n, L, m), e1 = digit_pos(M.n, L, m + 1);
        unsigned b = (unsigned)(e1 - e);
        int64_t acc = (int64_t)digits[m] + carry;
        int64_t low = acc & (int64_t)((1ULL << b) - 1);
        carry = (acc - low) >> b;
        or_bits(out, e, b, (uint64_t)low);
    }
    int64_t top = carry + xhi;
    if (top >= 0) {
        add_in(out, pow2n_times(M, (uint64_t)top));
    } else {
        uint64_t mag = (uint64_t)(-top);
        uint64_t j = (mag + M.k - 1) / M.k;      // j*k >= |top|
        uint64_t val = j * M.k - mag;            // 0 <= val < k
        add_in(out, pow2n_times(M, val));
        Limbs J(M.nl, 0);
        J[0] = j;
        add_in(out, J);
    }
    while (cmp(out, M.N) >= 0) sub_in(out, M.N);
    return out;
}

Limbs mul_kinv(const Modulus& M, const Limbs& x)
{
    Limbs y = mod_N(M, x);
    Limbs t = mul_small_mod(M, mul_2n(M, y), M.k - 1);  // x*(k-1)*2^n
    t.resize(M.nl + 1, 0);
    y.resize(M.nl + 1, 0);
    add_in(t, y);
    return mod_N(M, t);
}

Limbs kinv_pow(const Modulus& M, unsigned e)
{
    Limbs r(M.nl, 0);
    r[0] = 1;
    for (unsigned i = 0; i < e; ++i) r = mul_kinv(M, r);
    return r;
}

static uint64_t bitlen_limbs(const Limbs& x)
{
    for (size_t i = x.size(); i-- > 0;)
        if (x[i]) return 64 * (uint64_t)i + bitlen_u64(x[i]);
    return 0;
}

uint64_t distance_bits(const Modulus& M, const Limbs& x)
{
    Limbs y = mod_N(M, x);
    Limbs d = M.N;
    sub_in(d, y);  // N - y > 0
    return std::min(bitlen_limbs(y), bitlen_limbs(d));
}

uint64_t low64(const Limbs& x) { return x.empty() ? 0 : x[0]; }

bool is_n_minus_1(const Modulus& M, const Limbs& x)
{
    Limbs t = M.N;
    t[0] -= 1;  // N-1 = k*2^n (N[0] is 1 for n >= 64, else holds k<<n + 1)
    return cmp(x, t) == 0;
}

}  // namespace plbn
