// The FFT line lengths the box passes are compiled for: every 2/3/5-smooth
// length from 8 to 4096 (131 lengths), split into 8 groups so that the
// kernels of each group compile in a translation unit of their own
// (box_len_<g>.hip).  Each group holds one of the tuned lengths (512 ... 2560,
// those with alternative pass variants) and a share of the others.
#pragma once

#define SMK_FFT_LENGTHS_0(X) \
  X(16) X(45) X(75) X(120) X(216) X(288) X(360) X(512) X(675) X(720) X(1250) X(1350) X(1875) X(2700) X(2916) \
  X(4096)

#define SMK_FFT_LENGTHS_1(X) \
  X(20) X(30) X(96) X(100) X(240) X(270) X(405) X(576) X(640) X(810) X(1080) X(1600) X(2025) X(2400) X(3000) \
  X(4050)

#define SMK_FFT_LENGTHS_2(X) \
  X(25) X(32) X(90) X(108) X(225) X(256) X(432) X(540) X(768) X(864) X(1125) X(1500) X(2187) X(2250) X(3072) \
  X(4000)

#define SMK_FFT_LENGTHS_3(X) \
  X(9) X(18) X(50) X(60) X(150) X(162) X(320) X(384) X(648) X(729) X(1024) X(1215) X(1440) X(1920) X(2592) \
  X(3125) X(3888)

#define SMK_FFT_LENGTHS_4(X) \
  X(27) X(36) X(81) X(125) X(200) X(250) X(450) X(486) X(960) X(1000) X(1280) X(1620) X(2000) X(2430) X(3200) \
  X(3840)

#define SMK_FFT_LENGTHS_5(X) \
  X(8) X(15) X(54) X(64) X(135) X(192) X(300) X(400) X(625) X(750) X(1200) X(1296) X(1536) X(1800) X(2880) \
  X(3240) X(3750)

#define SMK_FFT_LENGTHS_6(X) \
  X(10) X(12) X(48) X(72) X(144) X(160) X(324) X(375) X(600) X(800) X(972) X(1728) X(1944) X(2048) X(2500) \
  X(3375) X(3645)

#define SMK_FFT_LENGTHS_7(X) \
  X(24) X(40) X(80) X(128) X(180) X(243) X(480) X(500) X(900) X(1152) X(1458) X(2160) X(2304) X(2560) X(3456) \
  X(3600)

#define SMK_FFT_LENGTHS(X)                                                                                       \
  SMK_FFT_LENGTHS_0(X) SMK_FFT_LENGTHS_1(X) SMK_FFT_LENGTHS_2(X) SMK_FFT_LENGTHS_3(X) SMK_FFT_LENGTHS_4(X) \
      SMK_FFT_LENGTHS_5(X) SMK_FFT_LENGTHS_6(X) SMK_FFT_LENGTHS_7(X)
