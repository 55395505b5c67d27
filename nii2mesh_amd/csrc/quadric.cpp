// quadric.cpp -- quadric_simplify_mesh (src/quadric.c:396-518) with its first
// update pass on the GPU.
//
// The reference simplifies by edge collapses in triangle order: a triangle
// collapses only if no earlier collapse of the same pass touched it, and each
// collapse reads what earlier collapses nearby have just changed.  That loop
// is sequential, so it runs here on the host, restated step for step.  Its
// most expensive order-free part -- update_mesh at iteration 0
// (src/quadric.c:144-215: reference lists, border flags, per-vertex plane
// quadrics, per-triangle edge errors) -- runs on the GPU (k_quadric.hip)
// with the reference's operation order, so the result is the reference's bit
// for bit.  Every FP64 expression below keeps the reference's order too.
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/nii2mesh_amd.h"
#include "n2m_internal.h"

using namespace n2m;

namespace {

struct QVert {
	vec3d p;
	int tstart, tcount;
	double q[10];
	int border;
};
struct QTri {
	int v[3];
	double err[4];
	bool dirty, deleted;
	vec3d n;
};
struct QRef {
	int tid, tvertex;
};

inline vec3d v3(double x, double y, double z) { return vec3d{x, y, z}; }
inline vec3d vsub(vec3d a, vec3d b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline vec3d vadd(vec3d a, vec3d b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline vec3d vscale(vec3d a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline double vdot(vec3d a, vec3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline vec3d vcross(vec3d a, vec3d b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline void vunit(vec3d *v) // src/quadric.c:87-93
{
	double len = sqrt((v->x * v->x) + (v->y * v->y) + (v->z * v->z));
	if (len <= 0)
		len = 0.001;
	v->x = v->x / len;
	v->y = v->y / len;
	v->z = v->z / len;
}

inline double qdet(const double *m, int a11, int a12, int a13, int a21, int a22, int a23, int a31, int a32, int a33)
{
	return m[a11] * m[a22] * m[a33] + m[a13] * m[a21] * m[a32] + m[a12] * m[a23] * m[a31] -
	       m[a13] * m[a22] * m[a31] - m[a11] * m[a23] * m[a32] - m[a12] * m[a21] * m[a33];
}

inline double qerr_at(const double *q, double x, double y, double z)
{
	return q[0] * x * x + 2 * q[1] * x * y + 2 * q[2] * x * z + 2 * q[3] * x + q[4] * y * y + 2 * q[5] * y * z +
	       2 * q[6] * y + q[7] * z * z + 2 * q[8] * z + q[9];
}

// calculate_error, src/quadric.c:108-133
double edge_error(int a, int b, vec3d *out, const QVert *V)
{
	double q[10];
	for (int i = 0; i < 10; i++)
		q[i] = V[a].q[i] + V[b].q[i];
	const int border = V[a].border + V[b].border;
	const double det = qdet(q, 0, 1, 2, 1, 4, 5, 2, 5, 7);
	if ((det != 0.0) && (border == 0)) {
		out->x = -1.0 / det * (qdet(q, 1, 2, 3, 4, 5, 6, 5, 7, 8));
		out->y = 1.0 / det * (qdet(q, 0, 2, 3, 1, 5, 6, 2, 7, 8));
		out->z = -1.0 / det * (qdet(q, 0, 1, 3, 1, 4, 6, 2, 5, 8));
		return qerr_at(q, out->x, out->y, out->z);
	}
	const vec3d p1 = V[a].p, p2 = V[b].p, p3 = vscale(vadd(p1, p2), 0.5);
	const double e1 = qerr_at(q, p1.x, p1.y, p1.z);
	const double e2 = qerr_at(q, p2.x, p2.y, p2.z);
	const double e3 = qerr_at(q, p3.x, p3.y, p3.z);
	const double e = fmin(e1, fmin(e2, e3));
	if (e1 == e)
		*out = p1;
	if (e2 == e)
		*out = p2;
	if (e3 == e)
		*out = p3;
	return e;
}

struct Mesh {
	std::vector<QVert> V;
	std::vector<QTri> T;
	std::vector<QRef> R; // reference lists, with room to grow (9 per triangle)
	int nref = 0, nT = 0, nV = 0;
	std::vector<char> del0, del1;
};

// update_mesh for iteration > 0 (src/quadric.c:144-178): drop deleted
// triangles (order kept), rebuild the reference lists in triangle order
void rebuild_refs(Mesh &M)
{
	int dst = 0;
	for (int i = 0; i < M.nT; i++)
		if (!M.T[i].deleted)
			M.T[dst++] = M.T[i];
	M.nT = dst;
	for (int i = 0; i < M.nV; i++) {
		M.V[i].tstart = 0;
		M.V[i].tcount = 0;
	}
	for (int i = 0; i < M.nT; i++)
		for (int j = 0; j < 3; j++)
			M.V[M.T[i].v[j]].tcount++;
	int ts = 0;
	for (int i = 0; i < M.nV; i++) {
		M.V[i].tstart = ts;
		ts += M.V[i].tcount;
		M.V[i].tcount = 0;
	}
	M.nref = M.nT * 3;
	for (int i = 0; i < M.nT; i++)
		for (int j = 0; j < 3; j++) {
			QVert &v = M.V[M.T[i].v[j]];
			M.R[v.tstart + v.tcount].tid = i;
			M.R[v.tstart + v.tcount].tvertex = j;
			v.tcount++;
		}
}

// update_mesh(iteration 0) from the GPU pass
void first_update(Mesh &M, const QuadricPrep &g)
{
	for (int i = 0; i < M.nV; i++) {
		QVert &v = M.V[i];
		v.tstart = (int)g.start[i];
		v.tcount = (int)(g.start[i + 1] - g.start[i]);
		v.border = g.border[i];
		memcpy(v.q, &g.q[10 * (size_t)i], sizeof v.q);
	}
	M.nref = M.nT * 3;
	for (int i = 0; i < M.nref; i++) {
		M.R[i].tid = (int)(g.ref[i] / 3);
		M.R[i].tvertex = (int)(g.ref[i] % 3);
	}
	for (int i = 0; i < M.nT; i++) {
		QTri &t = M.T[i];
		t.n = v3(g.nd[4 * (size_t)i], g.nd[4 * (size_t)i + 1], g.nd[4 * (size_t)i + 2]);
		for (int j = 0; j < 4; j++)
			t.err[j] = g.err[4 * (size_t)i + j];
	}
}

// flipped, src/quadric.c:295-314
bool flipped(vec3d p, int i1, const QVert &v0, char *del, Mesh &M)
{
	for (int k = 0; k < v0.tcount; k++) {
		const QRef &r = M.R[v0.tstart + k];
		const QTri &t = M.T[r.tid];
		if (t.deleted)
			continue;
		const int s = r.tvertex;
		const int id1 = t.v[(s + 1) % 3], id2 = t.v[(s + 2) % 3];
		if (id1 == i1 || id2 == i1) {
			del[k] = 1;
			continue;
		}
		vec3d d1 = vsub(M.V[id1].p, p);
		vunit(&d1);
		vec3d d2 = vsub(M.V[id2].p, p);
		vunit(&d2);
		if (fabs(vdot(d1, d2)) > 0.999)
			return true;
		vec3d n = vcross(d1, d2);
		vunit(&n);
		del[k] = 0;
		if (vdot(n, t.n) < 0.2)
			return true;
	}
	return false;
}

// update_triangles, src/quadric.c:272-293
void update_triangles(int i0, const QVert &v, const char *del, int *deleted, Mesh &M)
{
	vec3d p;
	for (int k = 0; k < v.tcount; k++) {
		const QRef r = M.R[v.tstart + k];
		QTri &t = M.T[r.tid];
		if (t.deleted)
			continue;
		if (del[k]) {
			t.deleted = 1;
			*deleted = *deleted + 1;
			continue;
		}
		t.v[r.tvertex] = i0;
		t.dirty = 1;
		t.err[0] = edge_error(t.v[0], t.v[1], &p, M.V.data());
		t.err[1] = edge_error(t.v[1], t.v[2], &p, M.V.data());
		t.err[2] = edge_error(t.v[2], t.v[0], &p, M.V.data());
		t.err[3] = fmin(t.err[0], fmin(t.err[1], t.err[2]));
		M.R[M.nref] = r;
		M.nref = M.nref + 1;
	}
}

// compact_mesh, src/quadric.c:246-270
void compact(Mesh &M, int *nT, int *nV)
{
	int dst = 0;
	for (int i = 0; i < *nV; i++)
		M.V[i].tcount = 0;
	for (int i = 0; i < *nT; i++)
		if (!M.T[i].deleted) {
			const QTri t = M.T[i];
			M.T[dst++] = t;
			for (int j = 0; j < 3; j++)
				M.V[t.v[j]].tcount = 1;
		}
	*nT = dst;
	dst = 0;
	for (int i = 0; i < *nV; i++)
		if (M.V[i].tcount) {
			M.V[i].tstart = dst;
			M.V[dst].p = M.V[i].p;
			dst++;
		}
	for (int i = 0; i < *nT; i++)
		for (int j = 0; j < 3; j++)
			M.T[i].v[j] = M.V[M.T[i].v[j]].tstart;
	*nV = dst;
}

} // namespace

extern "C" void quadric_simplify_mesh(vec3d **vs, vec3i **ts, int *nvert, int *ntri, int target_count,
                                      double agressiveness, bool verbose, bool finishLossless)
{
	Mesh M;
	const int nv0 = *nvert, nt0 = *ntri;
	QuadricPrep g;
	// the first update pass, on the GPU; nothing is changed when it fails
	if (quadric_prep(*vs, *ts, nv0, nt0, g) != 0) {
		fprintf(stderr, "nii2mesh_amd: quadric_simplify_mesh: %s\n", n2m_last_error());
		return;
	}
	M.V.resize(nv0 > 0 ? nv0 : 1);
	M.T.resize(nt0 > 0 ? nt0 : 1);
	M.R.resize((size_t)(nt0 > 0 ? nt0 : 1) * 9);
	M.del0.resize((size_t)(nt0 > 0 ? nt0 : 1) * 3);
	M.del1.resize((size_t)(nt0 > 0 ? nt0 : 1) * 3);
	for (int i = 0; i < nv0; i++)
		M.V[i].p = (*vs)[i];
	free(*vs);
	for (int i = 0; i < nt0; i++) {
		QTri &t = M.T[i];
		t.deleted = 0;
		t.dirty = 0;
		t.v[0] = (*ts)[i].x;
		t.v[1] = (*ts)[i].y;
		t.v[2] = (*ts)[i].z;
	}
	free(*ts);
	M.nV = nv0;
	M.nT = nt0;
	// src/quadric.c:425-491
	int deleted_triangles = 0;
	const int triangle_count = nt0;
	int max_iter = 100;
	bool lossy = true;
	double threshold = DBL_EPSILON;
	if (target_count >= M.nT) {
		lossy = false;
		max_iter = 1000;
	}
	int iterationStartCount = 0;
	for (int iteration = 0; iteration < max_iter; iteration++) {
		if ((lossy) && ((triangle_count - deleted_triangles) <= target_count)) {
			if (!finishLossless)
				break;
			lossy = false;
			threshold = DBL_EPSILON;
			max_iter = 1000;
		}
		if (lossy) {
			if (iteration % 5 == 0) {
				if (iteration == 0)
					first_update(M, g);
				else
					rebuild_refs(M);
			}
			threshold = 0.000000001 * pow((double)(iteration + 3.0), agressiveness);
		} else {
			if (iterationStartCount == (triangle_count - deleted_triangles))
				break;
			if (iteration == 0)
				first_update(M, g);
			else
				rebuild_refs(M);
		}
		iterationStartCount = triangle_count - deleted_triangles;
		for (int i = 0; i < M.nT; i++)
			M.T[i].dirty = 0;
		if ((verbose) && (iteration % 5 == 0))
			printf(" iteration %d - triangles %d threshold %g\n", iteration, triangle_count - deleted_triangles,
			       threshold);
		for (int i = 0; i < M.nT; i++) {
			QTri *t = &M.T[i];
			if (t->err[3] > threshold)
				continue;
			if (t->deleted)
				continue;
			if (t->dirty)
				continue;
			for (int j = 0; j < 3; j++)
				if (t->err[j] < threshold) {
					const int i0 = t->v[j];
					QVert *v0 = &M.V[i0];
					const int i1 = t->v[(j + 1) % 3];
					QVert *v1 = &M.V[i1];
					if (v0->border != v1->border)
						continue;
					vec3d p;
					edge_error(i0, i1, &p, M.V.data());
					if (flipped(p, i1, *v0, M.del0.data(), M))
						continue;
					if (flipped(p, i0, *v1, M.del1.data(), M))
						continue;
					v0->p = p;
					for (int k = 0; k < 10; k++) // symMatAdd(v0.q, v1.q, v0.q)
						v0->q[k] = v1->q[k] + v0->q[k];
					const int tstart = M.nref;
					update_triangles(i0, *v0, M.del0.data(), &deleted_triangles, M);
					update_triangles(i0, *v1, M.del1.data(), &deleted_triangles, M);
					const int tcount = M.nref - tstart;
					if (tcount <= v0->tcount) {
						if (tcount)
							memcpy(&M.R[v0->tstart], &M.R[tstart], tcount * sizeof(QRef));
					} else
						v0->tstart = tstart;
					v0->tcount = tcount;
					break;
				}
			if ((lossy) && ((triangle_count - deleted_triangles) <= target_count))
				break;
		}
	}
	int nT = M.nT, nV = M.nV;
	compact(M, &nT, &nV);
	*ntri = nT;
	*ts = (vec3i *)malloc((size_t)(nT > 0 ? nT : 1) * sizeof(vec3i));
	for (int i = 0; i < nT; i++)
		(*ts)[i] = vec3i{M.T[i].v[0], M.T[i].v[1], M.T[i].v[2]};
	*nvert = nV;
	*vs = (vec3d *)malloc((size_t)(nV > 0 ? nV : 1) * sizeof(vec3d));
	for (int i = 0; i < nV; i++)
		(*vs)[i] = M.V[i].p;
}
