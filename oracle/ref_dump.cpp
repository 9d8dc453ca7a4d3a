// Test infrastructure only: a tiny driver that links the reference's own Newell-tensor
// sources (Boris/DemagTFunc*.cpp, compiled from where they lie by oracle/Makefile) and
// writes the FP64 demag tensors it produces, so the oracle's restatement can be pinned
// bit-for-bit against them.  Nothing in the product links or loads this file.
//
// Usage:
//   ref_dump tensor3d nx ny nz Nx Ny Nz hx hy hz out.bin
//        -> Ddiag (Nx*Ny*Nz DBL3) then Dodiag (Nx*Ny*Nz DBL3), x fastest
//   ref_dump tensor2d nx ny Nx Ny hx hy hz out.bin
//        -> Ddiag (Nx*Ny DBL3) then Dodiag (Nx*Ny doubles)
//   ref_dump self hx hy hz out.bin      -> SelfDemag (3 doubles)
//   ref_dump single x y z hx hy hz out.bin -> Ldia_single (3) then Lodia_single (3)
//   ref_dump zshift2d nx ny Nx Ny sx sy sz dx dy dz shz out.bin
//        -> Ddiag then Dodiag (Nx*Ny DBL3 each, x fastest) of the 2D z-shifted tensor
//           between a source cell s and a destination cell d (both already normalised,
//           differing at most in z), shift shz (normalised): the irregular entry point,
//           which falls back to the regular shifted one when s == d
//   ref_dump pbc3d Nx Ny Nz hx hy hz px py pz out.bin
//        -> Ddiag then Dodiag (Nx*Ny*Nz DBL3 each) of the periodic tensor with px, py, pz
//           images (N = n along a periodic axis, 2n otherwise)
//   ref_dump selfpbc hx hy hz nx ny nz px py pz out.bin -> SelfDemag_PBC (3 doubles)
//   ref_dump pbc2d Nx Ny hx hy hz px py pz out.bin
//        -> Ddiag (Nx*Ny DBL3) then Dxy (Nx*Ny doubles) of the periodic 2D tensor
//   ref_dump shift2d nx ny Nx Ny sx sy sz dx dy dz shx shy shz out.bin
//        -> Ddiag then Dodiag (Nx*Ny DBL3 each) of the 2D tensor shifted in-plane and along z
//           (the general branch of the same entry point)
//   ref_dump zshift2dpbc Nx Ny sx sy sz dx dy dz shz px py out.bin
//        -> Ddiag then Dodiag (Nx*Ny DBL3 each) of the 2D z-shifted tensor with px, py
//           images in x and y (N = n along a periodic axis, 2n otherwise)
//
// Reference entry points used: DemagTFunc::CalcDiagTens3D/CalcOffDiagTens3D/
// CalcDiagTens2D/CalcOffDiagTens2D (Boris/DemagTFunc.cpp:21,95,172,237),
// DemagTFunc::SelfDemag (Boris/DemagTFunc_Self.cpp:46),
// DemagTFunc::Ldia_single/Lodia_single (Boris/DemagTFunc_Test.cpp),
// DemagTFunc::CalcDiagTens2D_Shifted_Irregular/CalcOffDiagTens2D_Shifted_Irregular
// (Boris/DemagTFunc_Irregular.cpp:5,121),
// DemagTFunc::CalcDiagTens3D_PBC/CalcOffDiagTens3D_PBC/CalcDiagTens2D_PBC/
// CalcOffDiagTens2D_PBC (Boris/DemagTFunc_PBC.cpp:186,301,5,96),
// DemagTFunc::SelfDemag_PBC (Boris/DemagTFunc_Self.cpp:58),
// DemagTFunc::CalcDiagTens2D_Shifted_Irregular_PBC/CalcOffDiagTens2D_Shifted_Irregular_PBC
// (Boris/DemagTFunc_Irregular_PBC.cpp:5, falling back to Boris/DemagTFunc_Shifted_PBC.cpp:317/464).

#include "stdafx.h"
#include "DemagTFunc.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static void write_doubles(const char* path, const std::vector<double>& v)
{
	FILE* f = fopen(path, "wb");
	if (!f) { perror(path); exit(2); }
	if (fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) { perror("fwrite"); exit(2); }
	fclose(f);
}

int main(int argc, char** argv)
{
	if (argc < 2) { fprintf(stderr, "usage: ref_dump tensor3d|tensor2d|self|single ...\n"); return 1; }
	DemagTFunc dtf;
	std::string mode = argv[1];

	if (mode == "tensor3d" && argc == 12) {
		INT3 n(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
		INT3 N(atoi(argv[5]), atoi(argv[6]), atoi(argv[7]));
		DBL3 h(atof(argv[8]), atof(argv[9]), atof(argv[10]));
		DBL3 hr = h / maximum(h.x, h.y, h.z);
		VEC<DBL3> Dd, Do;
		if (!Dd.resize(SZ3(N.x, N.y, N.z)) || !Do.resize(SZ3(N.x, N.y, N.z))) return 3;
		if (!dtf.CalcDiagTens3D(Dd, n, N, hr, true)) return 4;
		if (!dtf.CalcOffDiagTens3D(Do, n, N, hr)) return 4;
		std::vector<double> out;
		out.reserve((size_t)6 * Dd.linear_size());
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (size_t i = 0; i < Do.linear_size(); i++) { out.push_back(Do[i].x); out.push_back(Do[i].y); out.push_back(Do[i].z); }
		write_doubles(argv[11], out);
		return 0;
	}
	if (mode == "tensor2d" && argc == 10) {
		INT3 n(atoi(argv[2]), atoi(argv[3]), 1);
		INT3 N(atoi(argv[4]), atoi(argv[5]), 1);
		DBL3 h(atof(argv[6]), atof(argv[7]), atof(argv[8]));
		DBL3 hr = h / maximum(h.x, h.y, h.z);
		VEC<DBL3> Dd;
		if (!Dd.resize(SZ3(N.x, N.y, 1))) return 3;
		std::vector<double> Do((size_t)N.x * N.y, 0.0);
		if (!dtf.CalcDiagTens2D(Dd, n, N, hr, true)) return 4;
		if (!dtf.CalcOffDiagTens2D(Do, n, N, hr)) return 4;
		std::vector<double> out;
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (double d : Do) out.push_back(d);
		write_doubles(argv[9], out);
		return 0;
	}
	if (mode == "self" && argc == 6) {
		DBL3 h(atof(argv[2]), atof(argv[3]), atof(argv[4]));
		DBL3 s = dtf.SelfDemag(h, true);
		write_doubles(argv[5], { s.x, s.y, s.z });
		return 0;
	}
	if (mode == "single" && argc == 9) {
		DBL3 d(atof(argv[2]), atof(argv[3]), atof(argv[4]));
		DBL3 h(atof(argv[5]), atof(argv[6]), atof(argv[7]));
		DBL3 a = dtf.Ldia_single(d, h, true);
		DBL3 b = dtf.Lodia_single(d, h, true);
		write_doubles(argv[8], { a.x, a.y, a.z, b.x, b.y, b.z });
		return 0;
	}
	if (mode == "zshift2d" && argc == 14) {
		INT3 n(atoi(argv[2]), atoi(argv[3]), 1);
		INT3 N(atoi(argv[4]), atoi(argv[5]), 1);
		DBL3 s(atof(argv[6]), atof(argv[7]), atof(argv[8]));
		DBL3 d(atof(argv[9]), atof(argv[10]), atof(argv[11]));
		DBL3 shift(0.0, 0.0, atof(argv[12]));
		VEC<DBL3> Dd, Do;
		if (!Dd.resize(SZ3(N.x, N.y, 1)) || !Do.resize(SZ3(N.x, N.y, 1))) return 3;
		if (!dtf.CalcDiagTens2D_Shifted_Irregular(Dd, n, N, s, d, shift)) return 4;
		if (!dtf.CalcOffDiagTens2D_Shifted_Irregular(Do, n, N, s, d, shift)) return 4;
		std::vector<double> out;
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (size_t i = 0; i < Do.linear_size(); i++) { out.push_back(Do[i].x); out.push_back(Do[i].y); out.push_back(Do[i].z); }
		write_doubles(argv[13], out);
		return 0;
	}
	if (mode == "shift2d" && argc == 16) {
		INT3 n(atoi(argv[2]), atoi(argv[3]), 1);
		INT3 N(atoi(argv[4]), atoi(argv[5]), 1);
		DBL3 s(atof(argv[6]), atof(argv[7]), atof(argv[8]));
		DBL3 d(atof(argv[9]), atof(argv[10]), atof(argv[11]));
		DBL3 shift(atof(argv[12]), atof(argv[13]), atof(argv[14]));
		VEC<DBL3> Dd, Do;
		if (!Dd.resize(SZ3(N.x, N.y, 1)) || !Do.resize(SZ3(N.x, N.y, 1))) return 3;
		if (!dtf.CalcDiagTens2D_Shifted_Irregular(Dd, n, N, s, d, shift)) return 4;
		if (!dtf.CalcOffDiagTens2D_Shifted_Irregular(Do, n, N, s, d, shift)) return 4;
		std::vector<double> out;
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (size_t i = 0; i < Do.linear_size(); i++) { out.push_back(Do[i].x); out.push_back(Do[i].y); out.push_back(Do[i].z); }
		write_doubles(argv[15], out);
		return 0;
	}
	if (mode == "zshift2dpbc" && argc == 14) {
		INT3 N(atoi(argv[2]), atoi(argv[3]), 1);
		DBL3 s(atof(argv[4]), atof(argv[5]), atof(argv[6]));
		DBL3 d(atof(argv[7]), atof(argv[8]), atof(argv[9]));
		DBL3 shift(0.0, 0.0, atof(argv[10]));
		const int px = atoi(argv[11]), py = atoi(argv[12]);
		VEC<DBL3> Dd, Do;
		if (!Dd.resize(SZ3(N.x, N.y, 1)) || !Do.resize(SZ3(N.x, N.y, 1))) return 3;
		if (!dtf.CalcDiagTens2D_Shifted_Irregular_PBC(Dd, N, s, d, shift, true, ASYMPTOTIC_DISTANCE, px, py, 0)) return 4;
		if (!dtf.CalcOffDiagTens2D_Shifted_Irregular_PBC(Do, N, s, d, shift, true, ASYMPTOTIC_DISTANCE, px, py, 0)) return 4;
		std::vector<double> out;
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (size_t i = 0; i < Do.linear_size(); i++) { out.push_back(Do[i].x); out.push_back(Do[i].y); out.push_back(Do[i].z); }
		write_doubles(argv[13], out);
		return 0;
	}
	if (mode == "selfpbc" && argc == 12) {
		DBL3 h(atof(argv[2]), atof(argv[3]), atof(argv[4]));
		DBL3 n(atof(argv[5]), atof(argv[6]), atof(argv[7]));
		INT3 img(atoi(argv[8]), atoi(argv[9]), atoi(argv[10]));
		DBL3 s = dtf.SelfDemag_PBC(h, n, img);
		write_doubles(argv[11], { s.x, s.y, s.z });
		return 0;
	}
	if (mode == "pbc3d" && argc == 12) {
		INT3 N(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
		DBL3 h(atof(argv[5]), atof(argv[6]), atof(argv[7]));
		INT3 img(atoi(argv[8]), atoi(argv[9]), atoi(argv[10]));
		DBL3 hr = h / maximum(h.x, h.y, h.z);
		VEC<DBL3> Dd, Do;
		if (!Dd.resize(SZ3(N.x, N.y, N.z)) || !Do.resize(SZ3(N.x, N.y, N.z))) return 3;
		if (!dtf.CalcDiagTens3D_PBC(Dd, N, hr, true, ASYMPTOTIC_DISTANCE, img.x, img.y, img.z)) return 4;
		if (!dtf.CalcOffDiagTens3D_PBC(Do, N, hr, true, ASYMPTOTIC_DISTANCE, img.x, img.y, img.z)) return 4;
		std::vector<double> out;
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (size_t i = 0; i < Do.linear_size(); i++) { out.push_back(Do[i].x); out.push_back(Do[i].y); out.push_back(Do[i].z); }
		write_doubles(argv[11], out);
		return 0;
	}
	if (mode == "pbc2d" && argc == 11) {
		INT3 N(atoi(argv[2]), atoi(argv[3]), 1);
		DBL3 h(atof(argv[4]), atof(argv[5]), atof(argv[6]));
		INT3 img(atoi(argv[7]), atoi(argv[8]), atoi(argv[9]));
		DBL3 hr = h / maximum(h.x, h.y, h.z);
		VEC<DBL3> Dd;
		if (!Dd.resize(SZ3(N.x, N.y, 1))) return 3;
		std::vector<double> Do((size_t)N.x * N.y, 0.0);
		if (!dtf.CalcDiagTens2D_PBC(Dd, N, hr, true, ASYMPTOTIC_DISTANCE, img.x, img.y, img.z)) return 4;
		if (!dtf.CalcOffDiagTens2D_PBC(Do, N, hr, true, ASYMPTOTIC_DISTANCE, img.x, img.y, img.z)) return 4;
		std::vector<double> out;
		for (size_t i = 0; i < Dd.linear_size(); i++) { out.push_back(Dd[i].x); out.push_back(Dd[i].y); out.push_back(Dd[i].z); }
		for (double d : Do) out.push_back(d);
		write_doubles(argv[10], out);
		return 0;
	}
	fprintf(stderr, "bad arguments\n");
	return 1;
}
