// Generates hydra-lotus_amd/csrc/albedo_tables.inc: the directional albedo
// tables the glTF PBR closure reads for multiple-scattering compensation and
// layering (their roles: backend/vulkan/shaders/bsdf/albedo.slang:1-33).
//
//   g++ -O2 -std=c++17 -pthread make_albedo_tables.cpp -o make_albedo_tables
//   ./make_albedo_tables > ../csrc/albedo_tables.inc
//
// The layout is the reference's, so that a lookup means the same and the
// reference's committed table (backend/vulkan/bsdf-albedo-tables.inc) pins
// this one node for node: four tables, each a grid of nodes over
// coordinates in [0, 1] along the outgoing cosine, the alpha (the sheen's
// roughness) and, for the dielectric's three sheets, the relative index, on
// the axes of albedo.slang:78-236 (log(v) + k v, normalized; the leaving
// sheet's cosine axis split at the critical angle).
//
// Each node's albedo is the integral over the unit square of the lobe's
// sample weight f cos / p under its own sampler, which is the albedo by the
// change of variables:
//   E1 (a Fresnel of 1)          G2 / G1 of a visible normal;
//   Eschlick ((1 - cos_oh)^5)    (1 - cos_oh)^5 G2 / G1;
//   Edielectric (reflected)      F(cos_oh, eta) G2 / G1, a thin wall's
//                                2F / (1 + F);
//   Esheen                       2 pi f cos of a uniform hemisphere direction;
// a visible normal drawn by spherical caps (Dupuy and Benyoub 2023), a
// direction below the surface weighing 0. The integral is composite
// Gauss-Legendre (8 x 8 points a cell) from a 16 x 16 grid of cells, each
// cell split in four while its estimate and its quarters' sum differ by more
// than the cell's share of the tolerance, everything in double precision.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMinAlpha = 1.0e-3;       // preview.slang:17
constexpr double kMinSheenAxis = 0.005;    // albedo.slang:133
constexpr int kGgxCosines = 32, kDielectricCosines = 48, kAlphas = 32, kIndices = 16, kSheets = 3;
constexpr double kMinCosine = 2.0e-4, kCosineSlope = 4.0, kAlphaSlope = 2.0, kIndexSlope = 0.5;
constexpr double kMinIndex = 1.01, kMaxIndex = 3.0, kMinEvanescent = 2.0e-3;
constexpr int kCriticalNode = kDielectricCosines / 2;
constexpr double kTolerance = 2.0e-5;
constexpr int kMaxDepth = 7;

enum Table { kGgx = 0, kSchlick = 1, kDielectric = 2, kSheen = 3 };
enum Sheet { kEntering = 0, kLeaving = 1, kThin = 2 };

// --- axes ---------------------------------------------------------------------
double Scale(double v, double slope) { return std::log(v) + slope * v; }

// The value at coordinate x in [0, 1] of the axis from low to high: Newton's
// method on the concave scale, from below.
double LogLinearAt(double x, double low, double high, double slope) {
  const double target = Scale(low, slope) + (Scale(high, slope) - Scale(low, slope)) * std::clamp(x, 0.0, 1.0);
  double v = low;
  for (int i = 0; i < 200; ++i) v = std::clamp(v - (Scale(v, slope) - target) / (1.0 / v + slope), low, high);
  return v;
}
double CosineAt(double x) { return LogLinearAt(x, kMinCosine, 1.0, kCosineSlope); }
double AlphaAt(double x) { return LogLinearAt(x, kMinAlpha, 1.0, kAlphaSlope); }
double SheenAt(double x) { return LogLinearAt(x, kMinSheenAxis, 1.0, kAlphaSlope); }
double IndexAt(double x) { return 1.0 + LogLinearAt(x, kMinIndex - 1.0, kMaxIndex - 1.0, kIndexSlope); }

double CriticalCosine(double index) { return std::sqrt(std::max(1.0 - 1.0 / (index * index), 0.0)); }
double Evanescent(double mu, double index, double critical) {
  const double tau = std::max(index * std::sqrt(std::max(critical * critical - mu * mu, 0.0)), kMinEvanescent);
  return std::log(mu) - std::log(tau) + kCosineSlope * mu;
}

// The leaving sheet's cosine at coordinate x: below the critical node the
// evanescent scale by bisection, above it the refracted cosine's axis.
double LeavingAt(double x, double index) {
  const double xc = double(kCriticalNode) / double(kDielectricCosines - 1);
  const double critical = CriticalCosine(index);
  if (x <= xc) {
    const double low = Evanescent(kMinCosine, index, critical), high = Evanescent(critical, index, critical);
    const double target = low + (high - low) * (x / xc);
    double lo = kMinCosine, hi = critical;
    for (int i = 0; i < 200; ++i) {
      const double mid = 0.5 * (lo + hi);
      (Evanescent(mid, index, critical) < target ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
  }
  const double refracted = CosineAt((x - xc) / (1.0 - xc));
  return std::sqrt(1.0 - (1.0 - refracted * refracted) / (index * index));
}

// --- lobes ----------------------------------------------------------------------
struct V {
  double x, y, z;
};
V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double Dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V Normalize(V a) { return a * (1.0 / std::sqrt(Dot(a, a))); }

double GgxLambda(V w, double alpha) {
  const double tan2 = std::max(0.0, 1.0 - w.z * w.z) / (w.z * w.z);
  return 0.5 * (std::sqrt(1.0 + alpha * alpha * tan2) - 1.0);
}
V VisibleNormal(V wo, double alpha, double u1, double u2) {
  const V s = Normalize({wo.x * alpha, wo.y * alpha, wo.z});
  const double phi = 2.0 * kPi * u1;
  const double z = (1.0 - u2) * (1.0 + s.z) - s.z;
  const double st = std::sqrt(std::clamp(1.0 - z * z, 0.0, 1.0));
  const V h = V{st * std::cos(phi), st * std::sin(phi), z} + s;
  return Normalize({h.x * alpha, h.y * alpha, h.z});
}
double FresnelDielectric(double c, double eta) {
  c = std::clamp(c, 0.0, 1.0);
  const double s2 = (1.0 - c * c) / (eta * eta);
  if (s2 >= 1.0) return 1.0;
  const double ct = std::sqrt(1.0 - s2);
  const double par = (eta * c - ct) / (eta * c + ct), per = (c - eta * ct) / (c + eta * ct);
  return 0.5 * (par * par + per * per);
}

struct Node {
  int table;
  double mu, alpha, eta;
  bool thin;
};

// The sample weight at (u1, u2).
double Weight(const Node& n, double u1, double u2) {
  const V wo{std::sqrt(std::max(1.0 - n.mu * n.mu, 0.0)), 0.0, n.mu};
  if (n.table == kSheen) {
    const double st = std::sqrt(std::max(1.0 - u1 * u1, 0.0)), phi = 2.0 * kPi * u2;
    const V wi{st * std::cos(phi), st * std::sin(phi), u1};
    if (!(wi.z > 0.0)) return 0.0;
    const V h = Normalize(wo + wi);
    const double inverse = 1.0 / n.alpha;
    const double sin2 = std::max(1.0 - h.z * h.z, 0.0);
    const double d = (2.0 + inverse) * std::pow(sin2, 0.5 * inverse) / (2.0 * kPi);
    return 2.0 * kPi * d * wi.z / (4.0 * (wi.z + wo.z - wi.z * wo.z));
  }
  const V h = VisibleNormal(wo, n.alpha, u1, u2);
  const double c = Dot(wo, h);
  if (!(c > 0.0)) return 0.0;
  const V wi = h * (2.0 * c) + V{-wo.x, -wo.y, -wo.z};
  if (!(wi.z > 0.0)) return 0.0;
  const double lo = GgxLambda(wo, n.alpha), li = GgxLambda(wi, n.alpha);
  const double ratio = (1.0 + lo) / (1.0 + lo + li);
  double f = 1.0;
  if (n.table == kSchlick) {
    f = std::pow(std::clamp(1.0 - c, 0.0, 1.0), 5.0);
  } else if (n.table == kDielectric) {
    const double r = FresnelDielectric(c, n.eta);
    f = n.thin ? 2.0 * r / (1.0 + r) : r;
  }
  return f * ratio;
}

// 8-point Gauss-Legendre on [0, 1].
constexpr std::array<double, 8> kX = {0.01985507175123188, 0.10166676129318664, 0.2372337950418355,
                                      0.4082826787521751,  0.5917173212478249,  0.7627662049581645,
                                      0.8983332387068134,  0.9801449282487681};
constexpr std::array<double, 8> kW = {0.05061426814518813, 0.11119051722668724, 0.15685332293894363,
                                      0.18134189168918100, 0.18134189168918100, 0.15685332293894363,
                                      0.11119051722668724, 0.05061426814518813};

std::atomic<long long> g_evaluations{0};

double Cell(const Node& n, double x0, double y0, double size) {
  g_evaluations += 64;
  double s = 0.0;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) s += kW[i] * kW[j] * Weight(n, x0 + size * kX[i], y0 + size * kX[j]);
  return s * size * size;
}

double Adaptive(const Node& n, double x0, double y0, double size, double whole, double tolerance, int depth) {
  const double h = 0.5 * size;
  const double q[4] = {Cell(n, x0, y0, h), Cell(n, x0 + h, y0, h), Cell(n, x0, y0 + h, h),
                       Cell(n, x0 + h, y0 + h, h)};
  const double quarters = q[0] + q[1] + q[2] + q[3];
  if (depth >= kMaxDepth || std::fabs(quarters - whole) <= tolerance) return quarters;
  // a quarter's share: half the cell's, as a kink along a curve crosses
  // about twice as many cells at each level
  const double t = 0.5 * tolerance;
  return Adaptive(n, x0, y0, h, q[0], t, depth + 1) + Adaptive(n, x0 + h, y0, h, q[1], t, depth + 1) +
         Adaptive(n, x0, y0 + h, h, q[2], t, depth + 1) + Adaptive(n, x0 + h, y0 + h, h, q[3], t, depth + 1);
}

double Integrate(const Node& n) {
  // a 16 x 16 start grid (a coarser one can agree with its quarters by
  // chance near grazing directions), each cell refined on its own share of
  // the tolerance
  constexpr int kGrid = 16;
  constexpr double kSize = 1.0 / kGrid;
  double sum = 0.0;
  for (int i = 0; i < kGrid; ++i)
    for (int j = 0; j < kGrid; ++j) {
      const double x0 = kSize * i, y0 = kSize * j;
      sum += Adaptive(n, x0, y0, kSize, Cell(n, x0, y0, kSize), kTolerance / (kGrid * kGrid), 0);
    }
  return sum;
}

// The node's parameters, as albedo.slang:256-279 places them.
Node NodeAt(int table, int cosine, int alpha, int layer) {
  Node n{};
  n.table = table;
  const int cosines = table == kDielectric ? kDielectricCosines : kGgxCosines;
  const double xc = double(cosine) / double(cosines - 1), xa = double(alpha) / double(kAlphas - 1);
  if (table == kSheen) {
    n.mu = CosineAt(xc);
    n.alpha = SheenAt(xa);
    return n;
  }
  n.alpha = AlphaAt(xa);
  if (table != kDielectric) {
    n.mu = CosineAt(xc);
    return n;
  }
  const int sheet = layer / kIndices;
  const double index = IndexAt(double(layer % kIndices) / double(kIndices - 1));
  n.thin = sheet == kThin;
  if (sheet == kLeaving) {
    n.mu = LeavingAt(xc, index);
    n.eta = 1.0 / index;
  } else {
    n.mu = CosineAt(xc);
    n.eta = index;
  }
  return n;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 5) {  // one node: table cosine alpha layer, with its cost
    const Node n = NodeAt(std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    const double v = Integrate(n);
    std::printf("mu %.6g alpha %.6g eta %.6g thin %d: %.9g (%lld evaluations)\n", n.mu, n.alpha, n.eta, n.thin, v,
                static_cast<long long>(g_evaluations));
    return 0;
  }
  struct Slot {
    int table, cosine, alpha, layer;
  };
  std::vector<Slot> slots;
  for (int table = 0; table < 4; ++table) {
    const int cosines = table == kDielectric ? kDielectricCosines : kGgxCosines;
    const int layers = table == kDielectric ? kSheets * kIndices : 1;
    for (int layer = 0; layer < layers; ++layer)
      for (int a = 0; a < kAlphas; ++a)
        for (int c = 0; c < cosines; ++c) slots.push_back({table, c, a, layer});
  }
  std::vector<float> data(slots.size());
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  const unsigned threads = std::max(1u, std::thread::hardware_concurrency());
  for (unsigned t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (size_t i; (i = next.fetch_add(1)) < slots.size();) {
        const Slot& s = slots[i];
        data[i] = static_cast<float>(Integrate(NodeAt(s.table, s.cosine, s.alpha, s.layer)));
      }
    });
  for (auto& t : pool) t.join();
  std::printf("// Generated by hydra-lotus_amd/tables/make_albedo_tables.cpp: %zu floats, the\n"
              "// directional albedo tables E1, Eschlick, Edielectric (entering, leaving, thin)\n"
              "// and Esheen, cosine fastest, then alpha, then sheet and index.\n",
              data.size());
  for (size_t i = 0; i < data.size(); ++i) std::printf("%.9g,%s", data[i], (i + 1) % 8 == 0 ? "\n" : "");
  return 0;
}
