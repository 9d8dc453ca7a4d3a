// OVF 2.0 (OOMMF vector / scalar field file) reading and writing on the host, for
// magnetisation states and per-cell maps going in and out of the integrator.
//
// Follows OVF2::Read_OVF2_VEC / Read_OVF2_SCA / Write_OVF2_VEC / Write_OVF2_SCA
// (Boris/OVF2_Handlers.cpp:41-902): rectangular meshes only, meshunit "m" or "nm", valuedim
// 1 or 3, data as "binary 4" (led by the check value 1234567.0), "binary 8" (led by
// 123456789012345.0) or "text" (one cell per line, space or tab separated), x fastest.  The
// writer emits the reference's header lines in the reference's order with numbers printed
// at 6 significant digits (BorisLib ToString's default precision), so a file written here is
// byte-identical to the reference's for the same mesh, title and data.
#include "bm_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

namespace {

using bm::fail;

const char* const kHdrOvf2 = "# OOMMF OVF 2.0";
const char* const kMeshType = "# meshtype: ";
const char* const kMeshUnit = "# meshunit: ";
const char* const kValueDim = "# valuedim: ";
const char* const kMin[3] = { "# xmin: ", "# ymin: ", "# zmin: " };
const char* const kMax[3] = { "# xmax: ", "# ymax: ", "# zmax: " };
const char* const kNodes[3] = { "# xnodes: ", "# ynodes: ", "# znodes: " };
const char* const kStep[3] = { "# xstepsize: ", "# ystepsize: ", "# zstepsize: " };
const char* const kBeginData = "# Begin: data ";
const char* const kEndData = "# End: data ";

std::string lower(std::string s)
{
	for (char& c : s) c = (char)std::tolower((unsigned char)c);
	return s;
}

bool has(const std::string& line, const char* key) { return line.find(key) != std::string::npos; }
// the value after the key; the reference takes the substring after the key's length
std::string after(const std::string& line, const char* key) { return line.size() > strlen(key) ? line.substr(strlen(key)) : std::string(); }
double num(const std::string& s) { return std::strtod(s.c_str(), nullptr); }

// a text line without its trailing carriage return
bool read_line(std::istream& in, std::string& line)
{
	if (!std::getline(in, line)) return false;
	if (!line.empty() && line.back() == '\r') line.pop_back();
	return true;
}

struct Header {
	int n[3] = { 0, 0, 0 };
	double h[3] = { 0, 0, 0 };
	double s[3] = { 0, 0, 0 }, e[3] = { 0, 0, 0 };
	int valuedim = 0;
	int bytes = 0;          // 4, 8, or 1 for text
};

// Scans the header up to "# Begin: data ..." and leaves the stream at the data (after the
// check value for binary data).  Returns BM_OK or an error (reason in bm_last_error).
int scan(std::ifstream& in, Header& H)
{
	std::string line;
	if (!read_line(in, line) || line != kHdrOvf2) return fail(BM_ERR_ARG, "OVF2: not an OOMMF OVF 2.0 file");
	bool rectangular = false;
	double unit = 1.0;
	while (read_line(in, line)) {
		if (has(line, kMeshType)) {
			if (after(line, kMeshType) != "rectangular") return fail(BM_ERR_ARG, "OVF2: only rectangular meshes are read");
			rectangular = true;
		}
		else if (has(line, kMeshUnit)) {
			const std::string v = after(line, kMeshUnit);
			if (v == "m") unit = 1.0;
			else if (v == "nm") unit = 1e-9;
			else return fail(BM_ERR_ARG, "OVF2: meshunit must be m or nm");
		}
		else if (has(line, kValueDim)) {
			const std::string v = after(line, kValueDim);
			if (v == "1") H.valuedim = 1;
			else if (v == "3") H.valuedim = 3;
			else return fail(BM_ERR_ARG, "OVF2: valuedim must be 1 or 3");
		}
		else {
			bool hit = false;
			for (int a = 0; a < 3 && !hit; a++) {
				if (has(line, kMin[a])) { H.s[a] = num(after(line, kMin[a])) * unit; hit = true; }
				else if (has(line, kMax[a])) { H.e[a] = num(after(line, kMax[a])) * unit; hit = true; }
				else if (has(line, kNodes[a])) { H.n[a] = (int)num(after(line, kNodes[a])); hit = true; }
				else if (has(line, kStep[a])) { H.h[a] = num(after(line, kStep[a])) * unit; hit = true; }
			}
			if (hit) continue;
			if (lower(line).find(lower(kBeginData)) != std::string::npos) {
				const std::string v = lower(after(line, kBeginData));
				if (v == "binary 4") H.bytes = 4;
				else if (v == "binary 8") H.bytes = 8;
				else if (v == "text") H.bytes = 1;
				else continue;
				bool null_rect = true;
				for (int a = 0; a < 3; a++) null_rect = null_rect && H.s[a] == H.e[a];
				if (!rectangular || !H.valuedim || null_rect || !H.n[0] || !H.n[1] || !H.n[2] || !(H.h[0] > 0) || !(H.h[1] > 0) || !(H.h[2] > 0))
					return fail(BM_ERR_ARG, "OVF2: incomplete header");
				// the mesh the rectangle and cell size make must have the stated node counts
				for (int a = 0; a < 3; a++)
					if ((int)std::lround((H.e[a] - H.s[a]) / H.h[a]) != H.n[a]) return fail(BM_ERR_ARG, "OVF2: nodes do not match the rectangle and cell size");
				if (H.bytes == 4) {
					float c = 0;
					in.read(reinterpret_cast<char*>(&c), sizeof c);
					if (!in || c != 1234567.0f) return fail(BM_ERR_ARG, "OVF2: binary 4 check value missing");
				}
				else if (H.bytes == 8) {
					double c = 0;
					in.read(reinterpret_cast<char*>(&c), sizeof c);
					if (!in || c != 123456789012345.0) return fail(BM_ERR_ARG, "OVF2: binary 8 check value missing");
				}
				return BM_OK;
			}
		}
	}
	return fail(BM_ERR_ARG, "OVF2: no data section");
}

std::string fmt(double v)
{
	std::ostringstream ss;
	ss << std::setprecision(6) << v;
	return ss.str();
}

} // namespace

extern "C" int bm_ovf2_info(const char* path, int* n3, double* h3, double* rect6, int* valuedim, int* data_bytes)
{
	if (!path) return fail(BM_ERR_ARG, "bm_ovf2_info: null path");
	std::ifstream in(path, std::ios::in | std::ios::binary);
	if (!in.is_open()) return fail(BM_ERR_ARG, std::string("bm_ovf2_info: cannot open ") + path);
	Header H;
	int rc = scan(in, H);
	if (rc) return rc;
	for (int a = 0; a < 3; a++) {
		if (n3) n3[a] = H.n[a];
		if (h3) h3[a] = H.h[a];
		if (rect6) { rect6[a] = H.s[a]; rect6[3 + a] = H.e[a]; }
	}
	if (valuedim) *valuedim = H.valuedim;
	if (data_bytes) *data_bytes = H.bytes;
	return BM_OK;
}

extern "C" int bm_ovf2_read(const char* path, double* out, long long count)
{
	if (!path || !out) return fail(BM_ERR_ARG, "bm_ovf2_read: null argument");
	std::ifstream in(path, std::ios::in | std::ios::binary);
	if (!in.is_open()) return fail(BM_ERR_ARG, std::string("bm_ovf2_read: cannot open ") + path);
	Header H;
	int rc = scan(in, H);
	if (rc) return rc;
	const long long nv = (long long)H.n[0] * H.n[1] * H.n[2] * H.valuedim;
	if (count != nv) return fail(BM_ERR_ARG, "bm_ovf2_read: count must be valuedim * cells");
	if (H.bytes == 4) {
		std::vector<float> buf((size_t)nv);
		in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(nv * sizeof(float)));
		if (!in) return fail(BM_ERR_ARG, "bm_ovf2_read: file ends inside the data");
		for (long long q = 0; q < nv; q++) out[q] = buf[(size_t)q];
	}
	else if (H.bytes == 8) {
		in.read(reinterpret_cast<char*>(out), (std::streamsize)(nv * sizeof(double)));
		if (!in) return fail(BM_ERR_ARG, "bm_ovf2_read: file ends inside the data");
	}
	else {
		// one cell per line; a line with fewer fields than valuedim leaves the cell at zero,
		// as the reference leaves it unset
		std::string line;
		const long long ncell = nv / H.valuedim;
		for (long long c = 0; c < ncell; c++) {
			for (int q = 0; q < H.valuedim; q++) out[c * H.valuedim + q] = 0.0;
			if (!read_line(in, line)) return fail(BM_ERR_ARG, "bm_ovf2_read: file ends inside the text data");
			std::vector<double> f;
			size_t i = 0;
			while (i < line.size()) {
				while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) i++;
				size_t j = i;
				while (j < line.size() && line[j] != ' ' && line[j] != '\t') j++;
				if (j > i) f.push_back(num(line.substr(i, j - i)));
				i = j;
			}
			if ((int)f.size() >= H.valuedim)
				for (int q = 0; q < H.valuedim; q++) out[c * H.valuedim + q] = f[(size_t)q];
		}
	}
	return BM_OK;
}

extern "C" int bm_ovf2_write(const char* path, const double* data, int valuedim, const int* n3, const double* h3,
	const double* rect6, int data_type, double norm)
{
	if (!path || !data || !n3 || !h3 || !rect6) return fail(BM_ERR_ARG, "bm_ovf2_write: null argument");
	if (valuedim != 1 && valuedim != 3) return fail(BM_ERR_ARG, "bm_ovf2_write: valuedim must be 1 or 3");
	if (n3[0] < 1 || n3[1] < 1 || n3[2] < 1) return fail(BM_ERR_ARG, "bm_ovf2_write: bad mesh size");
	if (data_type != 0 && data_type != 4 && data_type != 8) return fail(BM_ERR_ARG, "bm_ovf2_write: data_type must be 4, 8 or 0 (text)");
	if (!(norm != 0)) return fail(BM_ERR_ARG, "bm_ovf2_write: norm must be non-zero");
	const char* dtype = data_type == 4 ? "binary 4" : data_type == 8 ? "binary 8" : "text";
	std::ofstream out(path, std::ios::out | std::ios::binary);
	if (!out.is_open()) return fail(BM_ERR_ARG, std::string("bm_ovf2_write: cannot create ") + path);
	// the title is the file name without its directory (ExtractFilenameDirectory)
	std::string title(path);
	const size_t slash = title.find_last_of("\\/");
	if (slash != std::string::npos) title = title.substr(slash + 1);
	out << kHdrOvf2 << "\n" << "#\n" << "# Segment count: 1\n" << "#\n" << "# Begin: Segment\n" << "# Begin: Header\n"
		<< "#\n" << "# Title: " << title << "\n" << "#\n" << kMeshUnit << "m\n" << "#\n" << kMeshType << "rectangular\n" << "#\n";
	for (int a = 0; a < 3; a++) out << kMin[a] << fmt(rect6[a]) << "\n";
	for (int a = 0; a < 3; a++) out << kMax[a] << fmt(rect6[3 + a]) << "\n";
	out << "#\n";
	for (int a = 0; a < 3; a++) out << kNodes[a] << n3[a] << "\n";
	out << "#\n";
	for (int a = 0; a < 3; a++) out << kStep[a] << fmt(h3[a]) << "\n";
	out << "#\n" << kValueDim << valuedim << "\n" << "#\n" << "# End: Header\n" << "#\n" << kBeginData << dtype << "\n";
	const long long nv = (long long)n3[0] * n3[1] * n3[2] * valuedim;
	if (data_type == 4) {
		const float c = 1234567.0f;
		out.write(reinterpret_cast<const char*>(&c), sizeof c);
		std::vector<float> buf((size_t)nv);
		for (long long q = 0; q < nv; q++) buf[(size_t)q] = (float)(data[q] / norm);
		out.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(nv * sizeof(float)));
	}
	else if (data_type == 8) {
		const double c = 123456789012345.0;
		out.write(reinterpret_cast<const char*>(&c), sizeof c);
		if (norm == 1.0) out.write(reinterpret_cast<const char*>(data), (std::streamsize)(nv * sizeof(double)));
		else {
			std::vector<double> buf((size_t)nv);
			for (long long q = 0; q < nv; q++) buf[(size_t)q] = data[q] / norm;
			out.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)(nv * sizeof(double)));
		}
	}
	else {
		const long long ncell = nv / valuedim;
		for (long long c = 0; c < ncell; c++) {
			for (int q = 0; q < valuedim; q++) out << (q ? " " : "") << data[c * valuedim + q] / norm;
			out << "\n";
		}
	}
	out << kEndData << dtype << "\n" << "# End: Segment\n";
	out.close();
	if (!out) return fail(BM_ERR_ARG, std::string("bm_ovf2_write: write failed for ") + path);
	return BM_OK;
}
