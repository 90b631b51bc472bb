// The library dequantisation tables: what DequantMatrices::Compute
// (pik/quant_weights.cc:607-926) produces for the predefined table set,
// i.e. what the reference decoder uses when the bitstream signals library
// tables (the reference encoder always does, pik/quant_weights.cc:928-940).
//
// Layout (the ABI's dequant_table / dequant_offsets): for each quant kind
// k = DCT8, ID, DCT2, DCT4, DCT16, DCT32, LINES and channel c, a run of
// 64, 64, 64, 64, 256, 1024, 64 floats starting at offsets[k * 3 + c].
//
// Values are numerator / weight in double, stored as float
// (ComputeQuantTable, pik/quant_weights.cc:416-566): the numerators are the
// IDCT scale products of the kind's transform, the weights come from the
// band parameters below through GetQuantWeights' distance / eccentricity
// interpolation (pik/quant_weights.cc:103-164). The parameters are the
// format's predefined library (pik/quant_weights.cc:617-878).
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "common.h"

namespace pikgpu {
namespace {

float V(double v) { return static_cast<float>(v); }

// IDCTScales<n> (pik/dct.h:25-130): iscale(u) = sqrt(2/n) eps(u)
// cos(u pi / 2n), eps(0) = sqrt(1/2); the AAN 8-point transform uses
// iscale, the recursive transforms of the other sizes 1 / (n iscale).
float IdctScale(int n, int u) {
  const double eps = u == 0 ? sqrt(0.5) : 1.0;
  const double iscale = sqrt(2.0 / n) * eps * cos(u * M_PI / (2.0 * n));
  return static_cast<float>(n == 8 ? iscale : 1.0 / (n * iscale));
}

double BandMult(double v) { return v > 0 ? 1 + v : 1 / (1 - v); }

double Interpolate(double pos, double max, const double* array, size_t len) {
  const double scaled_pos = pos * (len - 1) / max;
  const size_t idx = static_cast<size_t>(scaled_pos);
  const double a = array[idx];
  const double b = array[idx + 1];
  return a * pow(b / a, scaled_pos - idx);
}

// GetQuantWeights<sx, sy>: weight of coefficient (x, y) per channel from
// num_d distance bands and num_e eccentricity bands.
void BandWeights(int sx, int sy, const float* dist, int num_d,
                 const float* ecc, int num_e, double* out) {
  for (int c = 0; c < 3; ++c) {
    double bands[16];
    bands[0] = dist[c * num_d];
    for (int i = 1; i < num_d; ++i) {
      bands[i] = bands[i - 1] * BandMult(dist[c * num_d + i]);
    }
    double ebands[9];
    ebands[0] = 1.0;
    for (int i = 1; i <= num_e; ++i) {
      ebands[i] = ebands[i - 1] * BandMult(ecc[c * num_e + i - 1]);
    }
    for (int y = 0; y < sy; ++y) {
      for (int x = 0; x < sx; ++x) {
        const double dx = 1.0 * x / (sx - 1);
        const double dy = 1.0 * y / (sy - 1);
        const double distance = sqrt(dx * dx + dy * dy);
        const double wd = Interpolate(distance, sqrt(2) + 1e-6, bands, num_d);
        const double eccentricity =
            (x == 0 && y == 0) ? 0 : fabs(dx - dy) / distance;
        const double we =
            Interpolate(eccentricity, 1.0 + 1e-6, ebands, num_e + 1);
        out[c * sx * sy + y * sx + x] = we * wd;
      }
    }
  }
}

// Predefined library parameters.
const float kDct8Dist[3 * 6] = {
    V(6.7128322747011593),  V(-0.75596993600717899), V(-0.47741990264036249),
    V(-0.81596269409665323), V(0.068767170571484654), V(-20.887837229035178),
    V(1.0308008496910044),  V(0.12563824546332958),  V(-0.98580000474151119),
    V(-0.74783541528315123), V(-0.18837957703830949), V(-0.50540560621792985),
    V(0.52922318082990072), V(-6.2099138554436495),  V(2.4559360555622511),
    V(-1.8645272975104017), V(2.1626488944731781),   V(-20.514468231628619)};
const float kDct8Ecc[3 * 3] = {
    V(0.020372599856493687), V(0.0060672219272973112), V(-0.037634794641950318),
    V(0.19964780548254896),  V(-0.20598244512425934),  V(0.22606880802424917),
    V(-0.53357069165890425), V(0.067877070499761022),  V(2.3529080139232321)};
const float kIdWeights[3][3] = {
    {V(174.50360988711236), V(7098.4292418698387), V(4459.2881530953237)},
    {V(29.181414754407044), V(1462.9387613234978), V(1364.8889051351412)},
    {V(10.427519104606029), V(23.975682913740158), V(11.132318126587421)}};
const float kDct2Weights[3][6] = {
    {V(3838.4633860359086), V(2711.45620096628), V(740.86588368521473),
     V(673.9663156327548), V(146.0409913884842), V(71.829450601171018)},
    {V(855.89982430974862), V(835.22486787836522), V(268.7887798267422),
     V(161.58150295707284), V(46.818625352324425), V(28.025832307111365)},
    {V(135.95933746046285), V(100.36113442694905), V(52.759147600958094),
     V(54.55000110144173), V(10.61194822539392), V(6.7321557070577027)}};
const float kDct4Dist[3 * 4] = {
    V(20.464243458003235), V(-1.3216361675651374), V(-0.90068227414064506),
    V(-0.51692149442719293), V(3.4892753025959551), V(-0.3851659055605578),
    V(-1.6024424566582844), V(-0.090185175016963492), V(2.0543507462254667),
    V(-17.083007167897751), V(1.1553317008558754), V(-17.06851301189084)};
const float kDct4Ecc[3 * 2] = {V(-1.6540674814777321), V(1.4353603203078817),
                               V(0.23246389755392743), V(0.11670410074064763),
                               V(0.039676509798850998), V(1.7114284305197651)};
const float kDct4Muls[3][2] = {{V(0.47188805913083881), V(0.74665256923039514)},
                               {V(0.27688273718512119), V(0.32787026106006584)},
                               {V(0.94572969005995233), V(1.649348791638829)}};
const float kDct16Dist[3 * 6] = {
    V(2.8081053178832627),  V(-2.4300085829870786), V(0.11683860865233302),
    V(-0.48546810937737683), V(-772.68999845881376), V(-30.167218264433497),
    V(0.61651518963555374), V(-0.89670752611689697), V(-1.4823203833923126),
    V(-0.4392530120704895), V(-0.96459916681512592), V(-4.5043195385133448),
    V(0.35315014395417571), V(-6.1622959506013206), V(1.3987478239168303),
    V(-5.221619505420998),  V(-87.102308097158911), V(-29.330248661246706)};
const float kDct16Ecc[3 * 3] = {
    V(-0.1082223243760141), V(0.16581730095161393), V(-0.22834397719738264),
    V(0.064907061033690178), V(-0.07809582529363121), V(-0.044761862879806769),
    V(-0.23977989838080313), V(-0.14631104822608662), V(0.026626451443453436)};
const float kDct32Dist[3 * 8] = {
    V(0.84716094396432662), V(-2.4766455452218108), V(0.2471181572547147),
    V(0.57650543843415769), V(-4.0833701828342583), V(-28.279479541125081),
    V(1.8036899065163079),  V(39.052449003220673),
    V(0.17234631384979648), V(-1.1404450629580913), V(-0.69128963252295739),
    V(-0.53270455087075774), V(-0.46759485378919513), V(-0.89356535322414299),
    V(0.65008570941628885), V(-0.66302446211939114),
    V(0.22743363189568044), V(-11.670472775652776), V(-4.9179016084759626),
    V(-5.4264719484417459), V(-10.370646227045418), V(1.9002093523030437),
    V(-2.6705664701413623), V(-20.889766266401665)};
const float kDct32Ecc[3 * 4] = {
    V(-0.77613778421797797), V(0.8972017714545496), V(-0.93436764214829893),
    V(0.18670848590931757), V(0.089533427641859925), V(0.08358828409098),
    V(-0.094110728686133543), V(-0.1286652050040859), V(1.0095255806548),
    V(-1.5336522088790263), V(-6.9680701189357501), V(1.3664229471277314)};
const float kLinesPosition[64] = {
    0,   100, 100, 100, 100, 100, 100, 5, 100, 100, 50, 20, 20, 10, 5, 5,
    100, 100, 50,  20,  20,  10,  5,   5, 100, 50,  50, 20, 20, 10, 5, 5,
    100, 20,  20,  20,  20,  10,  5,   5, 100, 10,  10, 10, 10, 10, 5, 5,
    100, 5,   5,   5,   5,   5,   5,   5, 5,   5,   5,  5,  5,  5,  5, 5};
const float kLinesChannel[3] = {7.0f, 17.5f, 0.35f};
constexpr float kBad = 0xBAD;  // weight of positions that are never used

// Weights (3 x num) and numerators (num) of quant kind k.
int KindWeights(int k, double* w, double* num) {
  const int kSizes[kNumQuantKinds] = {64, 64, 64, 64, 256, 1024, 64};
  const int n = kSizes[k];
  switch (k) {
    case 0:   // DCT8
    case 4:   // DCT16
    case 5: {  // DCT32
      const int side = k == 0 ? 8 : (k == 4 ? 16 : 32);
      for (int i = 0; i < n; ++i) {
        const float s = IdctScale(side, i % side) * IdctScale(side, i / side);
        num[i] = s / n;
      }
      if (k == 0) BandWeights(8, 8, kDct8Dist, 6, kDct8Ecc, 3, w);
      if (k == 4) BandWeights(16, 16, kDct16Dist, 6, kDct16Ecc, 3, w);
      if (k == 5) BandWeights(32, 32, kDct32Dist, 8, kDct32Ecc, 4, w);
      break;
    }
    case 3: {  // DCT4: 4x4 weights, each covering a 2x2 of the interleaved
               // 8x8 layout, with extra multipliers on 01/10 and 11
      for (int i = 0; i < 64; ++i) {
        const float s = IdctScale(4, (i % 8) / 2) * IdctScale(4, (i / 8) / 2);
        num[i] = s / 16;
      }
      double w4[3 * 16];
      BandWeights(4, 4, kDct4Dist, 4, kDct4Ecc, 2, w4);
      for (int c = 0; c < 3; ++c) {
        for (int y = 0; y < 8; ++y) {
          for (int x = 0; x < 8; ++x) {
            w[c * 64 + y * 8 + x] = w4[c * 16 + (y / 2) * 4 + (x / 2)];
          }
        }
        w[c * 64 + 1] /= kDct4Muls[c][0];
        w[c * 64 + 8] /= kDct4Muls[c][0];
        w[c * 64 + 9] /= kDct4Muls[c][1];
      }
      break;
    }
    case 1: {  // identity
      for (int i = 0; i < 64; ++i) num[i] = 1.0;
      for (int c = 0; c < 3; ++c) {
        for (int i = 0; i < 64; ++i) w[64 * c + i] = kIdWeights[c][0];
        w[64 * c + 1] = kIdWeights[c][1];
        w[64 * c + 8] = kIdWeights[c][1];
        w[64 * c + 9] = kIdWeights[c][2];
      }
      break;
    }
    case 2: {  // DCT2X2: one weight per pyramid level and orientation
      for (int i = 0; i < 64; ++i) num[i] = 1.0;
      for (int c = 0; c < 3; ++c) {
        double* o = w + c * 64;
        const float* d = kDct2Weights[c];
        o[0] = kBad;
        o[1] = o[8] = d[0];
        o[9] = d[1];
        for (int y = 0; y < 2; ++y) {
          for (int x = 0; x < 2; ++x) {
            o[y * 8 + x + 2] = d[2];
            o[(y + 2) * 8 + x] = d[2];
            o[(y + 2) * 8 + x + 2] = d[3];
          }
        }
        for (int y = 0; y < 4; ++y) {
          for (int x = 0; x < 4; ++x) {
            o[y * 8 + x + 4] = d[4];
            o[(y + 4) * 8 + x] = d[4];
            o[(y + 4) * 8 + x + 4] = d[5];
          }
        }
      }
      break;
    }
    default: {  // LINES: scaled position weights (RawScaled)
      for (int i = 0; i < 64; ++i) num[i] = 1.0;
      for (int c = 0; c < 3; ++c) {
        for (int i = 0; i < 64; ++i) {
          const float base = i == 0 ? kBad : kLinesPosition[i];
          w[c * 64 + i] = base * kLinesChannel[c];
        }
      }
      break;
    }
  }
  return n;
}

}  // namespace

size_t DefaultDequantTables(float* table, size_t table_cap, uint32_t* offsets,
                            size_t offsets_cap, size_t* num_offsets) {
  double w[3 * 1024], num[1024];
  size_t pos = 0;
  for (int k = 0; k < kNumQuantKinds; ++k) {
    const int n = KindWeights(k, w, num);
    for (int c = 0; c < 3; ++c) {
      if (offsets && size_t(k * 3 + c) < offsets_cap) {
        offsets[k * 3 + c] = uint32_t(pos);
      }
      for (int i = 0; i < n; ++i, ++pos) {
        if (table && pos < table_cap) {
          table[pos] = static_cast<float>(num[i] / w[c * n + i]);
        }
      }
    }
  }
  if (num_offsets) *num_offsets = 3 * kNumQuantKinds;
  return pos;
}

}  // namespace pikgpu
