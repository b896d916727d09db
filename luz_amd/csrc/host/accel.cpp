// Acceleration structures and device packing.
//
// Topology parity: both builders reproduce the reference's build step by step in FP64
// and call the same libstdc++ algorithms with the same comparison outcomes
// (std::partition / std::nth_element for mesh BVHs, std::sort for the top level), so
// the node order, leaf ranges and triangle order are identical to the reference's.
//   mesh BVH (binned SAH, 12 bins, leaf <= 8)   Mesh.cpp:217-306,407-521
//   top-level BVH (median, n-ary leaves)         BVHNode.cpp:86-185
//   bounded/unbounded split                      Scene.cpp:655-688
//   light list + selection CDF                   Scene.cpp:615-638
//   camera                                       RayHelper.cpp:50-82
//   environment distribution                     EnvironmentMap.cpp:627-671
// The device then traverses a repacked form: binary 64-byte nodes holding both child
// boxes in fp32 (rounded outward) with leaves folded into their parent's reference.
#include "accel.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <functional>
#include <limits>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace luzgpu
{
	namespace
	{
		constexpr std::size_t	kMeshLeaf = 8;
		constexpr std::size_t	kBins = 12;
		constexpr std::size_t	kTopLeaf = 8;
		constexpr std::size_t	kTopRootLeaf = 16;

		AABBd	merge(const AABBd& a, const AABBd& b)
		{
			// Mesh.cpp:57-71 (std::min/std::max)
			AABBd r;
			for (int i = 0; i < 3; i++)
			{
				r.lo[i] = std::min(a.lo[i], b.lo[i]);
				r.hi[i] = std::max(a.hi[i], b.hi[i]);
			}
			return (r);
		}

		AABBd	mergeF(const AABBd& a, const AABBd& b)
		{
			// Utilities.cpp:58-70 (fmin/fmax)
			AABBd r;
			for (int i = 0; i < 3; i++)
			{
				r.lo[i] = std::fmin(a.lo[i], b.lo[i]);
				r.hi[i] = std::fmax(a.hi[i], b.hi[i]);
			}
			return (r);
		}

		double	surfaceArea(const AABBd& b)
		{
			const dvec3 e = b.hi - b.lo;
			const double x = std::max(0.0, e[0]), y = std::max(0.0, e[1]), z = std::max(0.0, e[2]);
			return (2.0 * ((x * y) + (x * z) + (y * z)));
		}

		int	largestAxis(const AABBd& b)
		{
			const dvec3 e = b.hi - b.lo;
			if (e[0] >= e[1] && e[0] >= e[2]) return (0);
			if (e[1] >= e[2]) return (1);
			return (2);
		}

		AABBd	triangleBox(const Triangle& t)
		{
			// Triangle.cpp:114-125: vertex bounds padded by T_MIN
			AABBd b;
			for (int i = 0; i < 3; i++)
			{
				b.lo[i] = std::min(t.v0[i], std::min(t.v1[i], t.v2[i])) - kTMin;
				b.hi[i] = std::max(t.v0[i], std::max(t.v1[i], t.v2[i])) + kTMin;
			}
			return (b);
		}

		struct	BinnedSplit
		{
			bool	found = false;
			int	axis = 0;
			std::size_t	splitBin = 0;
			double	centroidMin = 0.0;
			double	binScale = 0.0;
		};

		struct	Bin
		{
			AABBd	box;
			std::size_t	count = 0;
			bool	has = false;
		};

		std::size_t	binFor(double c, double mn, double scale)
		{
			std::size_t b = static_cast<std::size_t>((c - mn) * scale);
			return (b >= kBins ? kBins - 1 : b);
		}

		struct	BlasBuilder
		{
			std::vector<AABBd>	boxes;
			std::vector<dvec3>	centroids;
			Blas	out;

			BinnedSplit	findSplit(std::size_t start, std::size_t end)
			{
				const std::size_t n = end - start;
				const std::size_t minSide = std::max<std::size_t>(1, n / 16);
				BinnedSplit best;
				double bestCost = std::numeric_limits<double>::infinity();
				for (int axis = 0; axis < 3; axis++)
				{
					double cmin = std::numeric_limits<double>::infinity();
					double cmax = -std::numeric_limits<double>::infinity();
					for (std::size_t i = start; i < end; i++)
					{
						const double c = centroids[out.triangleIndices[i]][axis];
						cmin = std::min(cmin, c);
						cmax = std::max(cmax, c);
					}
					const double extent = cmax - cmin;
					if (extent <= 1e-12)
					{
						continue;
					}
					const double scale = static_cast<double>(kBins) / extent;
					std::array<Bin, kBins> bins;
					for (std::size_t i = start; i < end; i++)
					{
						const std::size_t t = out.triangleIndices[i];
						Bin& b = bins[binFor(centroids[t][axis], cmin, scale)];
						b.box = b.has ? merge(b.box, boxes[t]) : boxes[t];
						b.count++;
						b.has = true;
					}
					std::array<Bin, kBins> left, right;
					Bin run;
					for (std::size_t i = 0; i < kBins; i++)
					{
						if (bins[i].has)
						{
							run.box = run.has ? merge(run.box, bins[i].box) : bins[i].box;
							run.count += bins[i].count;
							run.has = true;
						}
						left[i] = run;
					}
					run = Bin();
					for (std::size_t i = kBins; i > 0; i--)
					{
						if (bins[i - 1].has)
						{
							run.box = run.has ? merge(run.box, bins[i - 1].box) : bins[i - 1].box;
							run.count += bins[i - 1].count;
							run.has = true;
						}
						right[i - 1] = run;
					}
					for (std::size_t s = 0; s + 1 < kBins; s++)
					{
						const Bin& l = left[s];
						const Bin& r = right[s + 1];
						if (!l.has || !r.has || l.count < minSide || r.count < minSide)
						{
							continue;
						}
						const double cost = (surfaceArea(l.box) * static_cast<double>(l.count))
							+ (surfaceArea(r.box) * static_cast<double>(r.count));
						if (cost < bestCost)
						{
							bestCost = cost;
							best = BinnedSplit{true, axis, s, cmin, scale};
						}
					}
				}
				return (best);
			}

			std::size_t	build(std::size_t start, std::size_t end)
			{
				const std::size_t index = out.nodes.size();
				const std::size_t n = end - start;
				AABBd box;
				for (std::size_t i = start; i < end; i++)
				{
					box = (i == start) ? boxes[out.triangleIndices[i]] : merge(box, boxes[out.triangleIndices[i]]);
				}
				out.nodes.push_back(BlasNode());
				out.nodes[index].box = box;
				if (n <= kMeshLeaf)
				{
					out.nodes[index].start = start;
					out.nodes[index].count = n;
					out.nodes[index].leaf = true;
					return (index);
				}
				const BinnedSplit split = findSplit(start, end);
				std::size_t mid = start;
				if (split.found)
				{
					auto it = std::partition(out.triangleIndices.begin() + start, out.triangleIndices.begin() + end,
						[&](std::size_t t) {
							return (binFor(centroids[t][split.axis], split.centroidMin, split.binScale) <= split.splitBin);
						});
					mid = static_cast<std::size_t>(it - out.triangleIndices.begin());
				}
				if (mid == start || mid == end)
				{
					const int axis = largestAxis(box);
					mid = start + n / 2;
					std::nth_element(out.triangleIndices.begin() + start, out.triangleIndices.begin() + mid,
						out.triangleIndices.begin() + end,
						[&](std::size_t a, std::size_t b) {
							const double ca = centroids[a][axis], cb = centroids[b][axis];
							if (ca == cb)
							{
								return (boxes[a].lo[axis] < boxes[b].lo[axis]);
							}
							return (ca < cb);
						});
				}
				const std::size_t l = build(start, mid);
				const std::size_t r = build(mid, end);
				out.nodes[index].left = l;
				out.nodes[index].right = r;
				return (index);
			}
		};

		// ---- top-level hittables ------------------------------------------------------

		struct	InstanceXf
		{
			dvec3	position, scale;
			double	cx = 1, sx = 0, cy = 1, sy = 0, cz = 1, sz = 0;

			explicit InstanceXf(const Instance& inst) : position(inst.position), scale(inst.scale)
			{
				// MeshInstance.cpp:48-75
				if (inst.hasTrig)
				{
					cx = inst.cosRotation[0]; sx = inst.sinRotation[0];
					cy = inst.cosRotation[1]; sy = inst.sinRotation[1];
					cz = inst.cosRotation[2]; sz = inst.sinRotation[2];
					return;
				}
				const double d2r = 3.14159265358979323846 / 180.0;
				cx = std::cos(inst.rotationDegrees[0] * d2r); sx = std::sin(inst.rotationDegrees[0] * d2r);
				cy = std::cos(inst.rotationDegrees[1] * d2r); sy = std::sin(inst.rotationDegrees[1] * d2r);
				cz = std::cos(inst.rotationDegrees[2] * d2r); sz = std::sin(inst.rotationDegrees[2] * d2r);
			}

			dvec3	rotate(dvec3 v) const
			{
				// MeshInstance.cpp:78-98
				double x = v[0], y = v[1], z = v[2];
				double ny = (y * cx) - (z * sx);
				double nz = (y * sx) + (z * cx);
				y = ny; z = nz;
				double nx = (x * cy) + (z * sy);
				nz = (-x * sy) + (z * cy);
				x = nx; z = nz;
				nx = (x * cz) - (y * sz);
				ny = (x * sz) + (y * cz);
				return {nx, ny, z};
			}

			dvec3	transformPoint(const dvec3& p) const
			{
				const dvec3 s = {p[0] * scale[0], p[1] * scale[1], p[2] * scale[2]};
				return (rotate(s) + position);
			}
		};

		bool	rectangleBasis(const Rectangle& r, dvec3& n, dvec3& wa, dvec3& ha)
		{
			// Rectangle.cpp:14-50 (no explicit axes)
			const dvec3& o = r.orientation;
			if (dot(o, o) <= 0.0)
			{
				return (false);
			}
			n = o * (1.0 / std::sqrt(dot(o, o)));
			if (std::fabs(n[1]) > 0.999) { wa = {1, 0, 0}; ha = {0, 0, 1}; return (true); }
			if (std::fabs(n[2]) > 0.999) { wa = {1, 0, 0}; ha = {0, 1, 0}; return (true); }
			if (std::fabs(n[0]) > 0.999) { wa = {0, 0, 1}; ha = {0, 1, 0}; return (true); }
			dvec3 helper = {0, 1, 0};
			if (std::fabs(dot(helper, n)) > 0.9)
			{
				helper = {1, 0, 0};
			}
			const dvec3 c1 = cross(helper, n);
			wa = c1 * (1.0 / std::sqrt(dot(c1, c1)));
			const dvec3 c2 = cross(n, wa);
			ha = c2 * (1.0 / std::sqrt(dot(c2, c2)));
			return (true);
		}

		bool	rectangleBox(const Rectangle& r, AABBd& out)
		{
			// Rectangle.cpp:295-343
			dvec3 n, wa, ha;
			if (r.width <= 0.0 || r.height <= 0.0 || !rectangleBasis(r, n, wa, ha))
			{
				return (false);
			}
			const double hw = r.width / 2.0, hh = r.height / 2.0;
			const dvec3 c0 = (r.position + (wa * hw)) + (ha * hh);
			const dvec3 c1 = (r.position + (wa * hw)) - (ha * hh);
			const dvec3 c2 = (r.position - (wa * hw)) + (ha * hh);
			const dvec3 c3 = (r.position - (wa * hw)) - (ha * hh);
			for (int i = 0; i < 3; i++)
			{
				out.lo[i] = std::min(std::min(c0[i], c1[i]), std::min(c2[i], c3[i]));
				out.hi[i] = std::max(std::max(c0[i], c1[i]), std::max(c2[i], c3[i]));
				if (std::fabs(out.hi[i] - out.lo[i]) <= kTMin)
				{
					out.lo[i] -= kTMin;
					out.hi[i] += kTMin;
				}
			}
			return (true);
		}

		GpuRect	packRect(const Rectangle& r)
		{
			dvec3 n{}, wa{}, ha{};
			rectangleBasis(r, n, wa, ha);
			GpuRect g{};
			for (int k = 0; k < 3; k++)
			{
				g.position[k] = static_cast<float>(r.position[k]);
				g.normal[k] = static_cast<float>(n[k]);
				g.widthAxis[k] = static_cast<float>(wa[k]);
				g.heightAxis[k] = static_cast<float>(ha[k]);
			}
			g.halfWidth = static_cast<float>(r.width / 2.0);
			g.halfHeight = static_cast<float>(r.height / 2.0);
			g.width = static_cast<float>(r.width);
			g.height = static_cast<float>(r.height);
			g.area = static_cast<float>(r.height * r.width);
			return (g);
		}

		// Cube::_generateFaces (Cube.cpp:72-95): front, back, top, bottom, right, left.
		std::array<Rectangle, 6>	cubeFaces(const Cube& c)
		{
			auto face = [](dvec3 pos, dvec3 o, double w, double h) {
				Rectangle r;
				r.position = pos;
				r.orientation = o;
				r.width = w;
				r.height = h;
				return r;
			};
			const dvec3 p = c.position;
			return {
				face(p - dvec3{0.0, 0.0, c.depth / 2.0}, {0.0, 0.0, -1.0}, c.width, c.height),
				face(p + dvec3{0.0, 0.0, c.depth / 2.0}, {0.0, 0.0, 1.0}, c.width, c.height),
				face(p + dvec3{0.0, c.height / 2.0, 0.0}, {0.0, 1.0, 0.0}, c.width, c.depth),
				face(p - dvec3{0.0, c.height / 2.0, 0.0}, {0.0, -1.0, 0.0}, c.width, c.depth),
				face(p + dvec3{c.width / 2.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, c.depth, c.height),
				face(p - dvec3{c.width / 2.0, 0.0, 0.0}, {-1.0, 0.0, 0.0}, c.depth, c.height),
			};
		}

		// Cube::createBoundingBox (Cube.cpp:166-183)
		AABBd	boxAround(const dvec3& p, double w, double h, double d)
		{
			AABBd b;
			b.lo = {p[0] - w / 2.0, p[1] - h / 2.0, p[2] - d / 2.0};
			b.hi = {p[0] + w / 2.0, p[1] + h / 2.0, p[2] + d / 2.0};
			return (b);
		}

		// Sphere::createBoundingBox (Sphere.cpp:352-360)
		AABBd	sphereBox(const dvec3& p, double r)
		{
			AABBd b;
			b.lo = p - dvec3{r, r, r};
			b.hi = p + dvec3{r, r, r};
			return (b);
		}

		bool	instanceBox(const Instance& inst, const Blas& blas, AABBd& out)
		{
			// MeshInstance.cpp:212-262
			if (blas.nodes.empty())
			{
				return (false);
			}
			const InstanceXf xf(inst);
			const AABBd& lb = blas.nodes[0].box;
			double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
			bool any = false;
			for (int c = 0; c < 8; c++)
			{
				const dvec3 corner = {(c & 1) ? lb.hi[0] : lb.lo[0], (c & 2) ? lb.hi[1] : lb.lo[1], (c & 4) ? lb.hi[2] : lb.lo[2]};
				const dvec3 w = xf.transformPoint(corner);
				if (!std::isfinite(w[0]) || !std::isfinite(w[1]) || !std::isfinite(w[2]))
				{
					continue;
				}
				for (int i = 0; i < 3; i++)
				{
					mn[i] = std::min(mn[i], w[i]);
					mx[i] = std::max(mx[i], w[i]);
				}
				any = true;
			}
			if (!any)
			{
				return (false);
			}
			out.lo = {mn[0], mn[1], mn[2]};
			out.hi = {mx[0], mx[1], mx[2]};
			return (true);
		}

		struct	TopNode
		{
			std::vector<std::int64_t>	children;	// >= 0 hittable, < 0: ~node index
			std::vector<AABBd>	childBoxes;
			AABBd	box;
		};

		struct	TlasBuilder
		{
			const std::vector<AABBd>&	boxes;	// per scene hittable
			std::vector<TopNode>	nodes;

			std::int64_t	build(std::vector<std::uint32_t>& items, std::size_t start, std::size_t end, bool root)
			{
				// BVHNode.cpp:142-185
				AABBd range;
				for (std::size_t i = start; i < end; i++)
				{
					range = (i == start) ? boxes[items[i]] : mergeF(range, boxes[items[i]]);
				}
				const int axis = largestAxis(range);
				const std::size_t n = end - start;
				const std::size_t self = nodes.size();
				nodes.push_back(TopNode());
				TopNode node;
				if (n <= kTopLeaf || (root && n <= kTopRootLeaf))
				{
					for (std::size_t i = start; i < end; i++)
					{
						node.children.push_back(items[i]);
						node.childBoxes.push_back(boxes[items[i]]);
					}
				}
				else
				{
					std::sort(items.begin() + start, items.begin() + end, [&](std::uint32_t a, std::uint32_t b) {
						return (boxes[a].lo[axis] < boxes[b].lo[axis]);
					});
					const std::size_t mid = start + n / 2;
					const std::int64_t l = build(items, start, mid, false);
					const std::int64_t r = build(items, mid, end, false);
					node.children = {~l, ~r};
					node.childBoxes = {nodes[static_cast<std::size_t>(l)].box, nodes[static_cast<std::size_t>(r)].box};
				}
				for (std::size_t i = 0; i < node.childBoxes.size(); i++)
				{
					node.box = (i == 0) ? node.childBoxes[0] : mergeF(node.box, node.childBoxes[i]);
				}
				nodes[self] = node;
				return (static_cast<std::int64_t>(self));
			}

			void	dump(std::ostringstream& os, std::int64_t node) const
			{
				const TopNode& n = nodes[static_cast<std::size_t>(node)];
				os << "(";
				for (std::size_t i = 0; i < n.children.size(); i++)
				{
					if (i) os << ",";
					if (n.children[i] >= 0) os << n.children[i];
					else dump(os, ~n.children[i]);
				}
				os << ")";
			}
		};

		float	roundDown(double v)
		{
			float f = static_cast<float>(v);
			if (static_cast<double>(f) > v) f = std::nextafter(f, -INFINITY);
			return (f);
		}

		float	roundUp(double v)
		{
			float f = static_cast<float>(v);
			if (static_cast<double>(f) < v) f = std::nextafter(f, INFINITY);
			return (f);
		}

		void	setChild(BvhNode& node, int slot, const AABBd* box, std::int32_t child, std::int32_t count)
		{
			float* lo = slot == 0 ? node.lo0 : node.lo1;
			float* hi = slot == 0 ? node.hi0 : node.hi1;
			for (int i = 0; i < 3; i++)
			{
				if (box)
				{
					lo[i] = roundDown(box->lo[i]);
					hi[i] = roundUp(box->hi[i]);
				}
				else
				{
					lo[i] = INFINITY;	// empty slot: no ray enters it
					hi[i] = -INFINITY;
				}
			}
			node.child[slot] = child;
			node.count[slot] = count;
		}

		int	deviceDepth(const std::vector<BvhNode>& nodes, std::int32_t root)
		{
			if (root < 0) return 0;
			const BvhNode& n = nodes[static_cast<std::size_t>(root)];
			int d = 0;
			for (int s = 0; s < 2; s++)
			{
				if (n.count[s] == 0) d = std::max(d, deviceDepth(nodes, n.child[s]));
			}
			return 1 + d;
		}

		struct	TlasEmitter
		{
			const std::vector<AABBd>&	boxes;
			const std::vector<GpuHittable>&	hittables;
			std::vector<BvhNode>&	nodes;
			std::vector<std::uint32_t>&	prims;

			struct	Ref
			{
				std::int32_t	child;
				std::int32_t	count;	// 0: node, > 0: primitive range
				AABBd	box;
			};

			Ref	group(std::vector<std::uint32_t> items)
			{
				AABBd box = boxes[items[0]];
				bool instances = false;
				for (std::uint32_t h : items)
				{
					box = merge(box, boxes[h]);
					instances = instances || (hittables[h].kindIndex >> kKindShift) == GPU_HIT_INSTANCE;
				}
				if (items.size() == 1 || (!instances && items.size() <= 4))
				{
					const std::int32_t start = static_cast<std::int32_t>(prims.size());
					prims.insert(prims.end(), items.begin(), items.end());
					return Ref{start, static_cast<std::int32_t>(items.size()), box};
				}
				AABBd cb;
				for (std::size_t i = 0; i < items.size(); i++)
				{
					const dvec3 c = (boxes[items[i]].lo + boxes[items[i]].hi) * 0.5;
					const AABBd cc{c, c};
					cb = i == 0 ? cc : merge(cb, cc);
				}
				const int axis = largestAxis(cb);
				const std::size_t mid = items.size() / 2;
				std::nth_element(items.begin(), items.begin() + static_cast<std::ptrdiff_t>(mid), items.end(),
					[&](std::uint32_t a, std::uint32_t b) {
						const double ca = boxes[a].lo[axis] + boxes[a].hi[axis], cbv = boxes[b].lo[axis] + boxes[b].hi[axis];
						return ca < cbv || (ca == cbv && a < b);
					});
				const std::size_t idx = nodes.size();
				nodes.push_back(BvhNode{});
				const Ref l = group(std::vector<std::uint32_t>(items.begin(), items.begin() + static_cast<std::ptrdiff_t>(mid)));
				const Ref r = group(std::vector<std::uint32_t>(items.begin() + static_cast<std::ptrdiff_t>(mid), items.end()));
				setChild(nodes[idx], 0, &l.box, l.child, l.count);
				setChild(nodes[idx], 1, &r.box, r.child, r.count);
				return Ref{static_cast<std::int32_t>(idx), 0, box};
			}

			Ref	fromReference(const TlasBuilder& tb, std::size_t node)
			{
				const TopNode& n = tb.nodes[node];
				if (n.children.empty() || n.children[0] >= 0)
				{
					std::vector<std::uint32_t> items;
					for (std::int64_t c : n.children) items.push_back(static_cast<std::uint32_t>(c));
					Ref r = group(items);
					r.box = n.box;
					return r;
				}
				const std::size_t idx = nodes.size();
				nodes.push_back(BvhNode{});
				const Ref l = fromReference(tb, static_cast<std::size_t>(~n.children[0]));
				const Ref r = fromReference(tb, static_cast<std::size_t>(~n.children[1]));
				setChild(nodes[idx], 0, &l.box, l.child, l.count);
				setChild(nodes[idx], 1, &r.box, r.child, r.count);
				return Ref{static_cast<std::int32_t>(idx), 0, n.box};
			}
		};

		// ---- 4-wide device nodes --------------------------------------------------------------
		struct	WideSlot
		{
			float	lo[3], hi[3];
			std::int32_t	child;
			std::int32_t	count;	// 0: binary node index (internal), > 0 leaf range
		};

		float	slotArea(const WideSlot& s)
		{
			const float x = s.hi[0] - s.lo[0], y = s.hi[1] - s.lo[1], z = s.hi[2] - s.lo[2];
			return x * y + x * z + y * z;
		}

		void	encodeWide(Bvh4Node& n, const std::vector<WideSlot>& slots)
		{
			float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
			for (const WideSlot& s : slots)
				for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], s.lo[a]); hi[a] = std::max(hi[a], s.hi[a]); }
			for (int a = 0; a < 3; a++)
			{
				n.origin[a] = lo[a];
				const double extent = static_cast<double>(hi[a]) - static_cast<double>(lo[a]);
				int e = extent > 0.0 ? static_cast<int>(std::ceil(std::log2(extent / 255.0))) : -100;
				e = std::max(e, -100);
				for (;; e++)
				{
					if (e > 127) throw std::runtime_error("BVH box too large to quantise");
					const float scale = std::ldexp(1.0f, e);
					bool ok = true;
					for (std::size_t k = 0; k < slots.size() && ok; k++)
					{
						long ql = static_cast<long>(std::floor((static_cast<double>(slots[k].lo[a]) - lo[a]) / scale));
						long qh = static_cast<long>(std::ceil((static_cast<double>(slots[k].hi[a]) - lo[a]) / scale));
						ql = std::max(0L, std::min(255L, ql));
						qh = std::max(0L, std::min(255L, qh));
						while (ql > 0 && std::fma(static_cast<float>(ql), scale, lo[a]) > slots[k].lo[a]) ql--;
						while (qh < 255 && std::fma(static_cast<float>(qh), scale, lo[a]) < slots[k].hi[a]) qh++;
						if (std::fma(static_cast<float>(ql), scale, lo[a]) > slots[k].lo[a]
							|| std::fma(static_cast<float>(qh), scale, lo[a]) < slots[k].hi[a])
						{
							ok = false;
							break;
						}
						n.qlo[a][k] = static_cast<std::uint8_t>(ql);
						n.qhi[a][k] = static_cast<std::uint8_t>(qh);
					}
					if (ok)
					{
						n.exp[a] = static_cast<std::int8_t>(e);
						break;
					}
				}
			}
			for (std::size_t k = 0; k < 4; k++)
			{
				if (k < slots.size())
				{
					n.child[k] = slots[k].child;
					n.count[k] = static_cast<std::uint8_t>(slots[k].count);
				}
				else
				{
					n.child[k] = 0;
					n.count[k] = 255;
					for (int a = 0; a < 3; a++) { n.qlo[a][k] = 255; n.qhi[a][k] = 0; }
				}
			}
		}

		// Collapses the binary device tree below `node` into 4-wide nodes (largest-area
		// internal child opened first). Returns the wide index of the new node.
		std::int32_t	widen(const std::vector<BvhNode>& bin, std::int32_t node, std::vector<Bvh4Node>& out)
		{
			std::vector<WideSlot> slots;
			auto addChildren = [&](std::int32_t b) {
				const BvhNode& n = bin[static_cast<std::size_t>(b)];
				for (int s = 0; s < 2; s++)
				{
					if (n.count[s] < 0) continue;
					WideSlot w;
					const float* lo = s == 0 ? n.lo0 : n.lo1;
					const float* hi = s == 0 ? n.hi0 : n.hi1;
					for (int a = 0; a < 3; a++) { w.lo[a] = lo[a]; w.hi[a] = hi[a]; }
					w.child = n.child[s];
					w.count = n.count[s];
					slots.push_back(w);
				}
			};
			addChildren(node);
			while (slots.size() < 4)
			{
				int best = -1;
				float bestArea = -1.0f;
				for (std::size_t k = 0; k < slots.size(); k++)
				{
					if (slots[k].count == 0 && slotArea(slots[k]) > bestArea)
					{
						best = static_cast<int>(k);
						bestArea = slotArea(slots[k]);
					}
				}
				if (best < 0) break;
				const BvhNode& open = bin[static_cast<std::size_t>(slots[static_cast<std::size_t>(best)].child)];
				const int kids = (open.count[0] >= 0 ? 1 : 0) + (open.count[1] >= 0 ? 1 : 0);
				if (slots.size() - 1 + static_cast<std::size_t>(kids) > 4) break;
				const std::int32_t b = slots[static_cast<std::size_t>(best)].child;
				slots.erase(slots.begin() + best);
				addChildren(b);
			}
			const std::size_t idx = out.size();
			out.push_back(Bvh4Node{});
			for (WideSlot& w : slots)
			{
				if (w.count == 0) w.child = widen(bin, w.child, out);
			}
			encodeWide(out[idx], slots);
			return static_cast<std::int32_t>(idx);
		}

		// Worst-case traversal stack entries below `root`: a node defers every hit child but
		// the nearest, so the bound is the largest sum of (children - 1) along a root-to-leaf path.
		int	wideStackBound(const std::vector<Bvh4Node>& nodes, std::int32_t root)
		{
			const Bvh4Node& n = nodes[static_cast<std::size_t>(root)];
			int valid = 0, below = 0;
			for (int k = 0; k < 4; k++)
			{
				if (n.count[k] == 255) continue;
				valid++;
				if (n.count[k] == 0) below = std::max(below, wideStackBound(nodes, n.child[k]));
			}
			return std::max(0, valid - 1) + below;
		}

		int	wideDepth(const std::vector<Bvh4Node>& nodes, std::int32_t root)
		{
			const Bvh4Node& n = nodes[static_cast<std::size_t>(root)];
			int d = 0;
			for (int k = 0; k < 4; k++)
			{
				if (n.count[k] == 0) d = std::max(d, wideDepth(nodes, n.child[k]));
			}
			return 1 + d;
		}

		GpuTri	packTriangle(const Triangle& t, std::int32_t prim)
		{
			GpuTri g{};
			const dvec3 e1 = t.v1 - t.v0, e2 = t.v2 - t.v0;
			for (int i = 0; i < 3; i++)
			{
				g.v0[i] = static_cast<float>(t.v0[i]);
				g.e1[i] = static_cast<float>(e1[i]);
				g.e2[i] = static_cast<float>(e2[i]);
			}
			g.prim = prim;
			return (g);
		}

		// Mesh BVH -> binary device nodes appended to `nodes`; returns the root index.
		std::int32_t	packBlas(const Blas& blas, std::int32_t triBase, std::vector<BvhNode>& nodes)
		{
			std::vector<std::int32_t> map(blas.nodes.size(), -1);
			const std::int32_t base = static_cast<std::int32_t>(nodes.size());
			// device nodes for internal reference nodes, in reference (preorder) order
			std::int32_t next = base;
			for (std::size_t i = 0; i < blas.nodes.size(); i++)
			{
				if (!blas.nodes[i].leaf)
				{
					map[i] = next++;
				}
			}
			const bool rootLeaf = blas.nodes.empty() || blas.nodes[0].leaf;
			nodes.resize(static_cast<std::size_t>(next) + (rootLeaf ? 1 : 0));
			if (rootLeaf)
			{
				BvhNode& r = nodes[static_cast<std::size_t>(next)];
				if (blas.nodes.empty())
				{
					setChild(r, 0, nullptr, 0, -1);
				}
				else
				{
					setChild(r, 0, &blas.nodes[0].box, triBase + static_cast<std::int32_t>(blas.nodes[0].start),
						static_cast<std::int32_t>(blas.nodes[0].count));
				}
				setChild(r, 1, nullptr, 0, -1);
				return (next);
			}
			for (std::size_t i = 0; i < blas.nodes.size(); i++)
			{
				const BlasNode& n = blas.nodes[i];
				if (n.leaf)
				{
					continue;
				}
				BvhNode& d = nodes[static_cast<std::size_t>(map[i])];
				const std::size_t kids[2] = {n.left, n.right};
				for (int s = 0; s < 2; s++)
				{
					const BlasNode& c = blas.nodes[kids[s]];
					if (c.leaf)
					{
						setChild(d, s, &c.box, triBase + static_cast<std::int32_t>(c.start), static_cast<std::int32_t>(c.count));
					}
					else
					{
						setChild(d, s, &c.box, map[kids[s]], 0);
					}
				}
			}
			return (map[0]);
		}

		dvec3	normalizeD(const dvec3& v)
		{
			return (v * (1.0 / std::sqrt(dot(v, v))));
		}
	}

	Blas	buildBlas(const Mesh& mesh)
	{
		BlasBuilder b;
		const std::size_t n = mesh.triangles.size();
		if (n == 0)
		{
			return (Blas());
		}
		b.boxes.reserve(n);
		b.centroids.reserve(n);
		b.out.triangleIndices.resize(n);
		std::iota(b.out.triangleIndices.begin(), b.out.triangleIndices.end(), 0);
		for (const Triangle& t : mesh.triangles)
		{
			const AABBd box = triangleBox(t);
			b.boxes.push_back(box);
			// boxCentroid: (min + max) / 2.0 == 0.5 * (min + max)
			b.centroids.push_back((box.lo + box.hi) * 0.5);
		}
		b.out.nodes.reserve(2 * ((n + kMeshLeaf - 1) / kMeshLeaf));
		b.build(0, n);
		return (std::move(b.out));
	}

	GpuCamera	makeCamera(const Camera& cam, std::uint32_t width, std::uint32_t height, double metersPerUnit)
	{
		// RayHelper.cpp:50-82
		GpuCamera gc{};
		const double W = width, H = height;
		gc.inverseWidthMinusOne = W > 1.0 ? 1.0 / (W - 1.0) : 0.0;
		gc.inverseHeightMinusOne = H > 1.0 ? 1.0 / (H - 1.0) : 0.0;
		const double focus = cam.focusDistance / metersPerUnit;
		double vw = focus * cam.sensorWidth / cam.focalLength;
		double vh = focus * cam.sensorHeight / cam.focalLength;
		const double imageAspect = W / H;
		const double viewportAspect = vw / vh;
		if (imageAspect > viewportAspect) vh = vw / imageAspect;
		else if (imageAspect < viewportAspect) vw = vh * imageAspect;
		const dvec3 w = normalizeD(cam.direction);
		dvec3 up = cam.up;
		if (!std::isfinite(up[0]) || !std::isfinite(up[1]) || !std::isfinite(up[2]) || dot(up, up) <= 1e-12)
		{
			up = {0.0, 1.0, 0.0};
		}
		up = normalizeD(up);
		if (dot(cross(up, w), cross(up, w)) <= 1e-12) up = {0.0, 0.0, 1.0};
		if (dot(cross(up, w), cross(up, w)) <= 1e-12) up = {1.0, 0.0, 0.0};
		const dvec3 u = normalizeD(cross(up, w));
		const dvec3 v = cross(w, u);
		const double aperture = cam.pinhole ? 0.0 : cam.focalLength / cam.fNumber;
		gc.lensRadius = aperture / (2.0 * metersPerUnit);
		const dvec3 horizontal = u * vw;
		const dvec3 vertical = v * vh;
		const dvec3 llc = ((cam.position + (horizontal * (1.0 / 2.0))) + (vertical * (1.0 / 2.0))) + (w * focus);
		for (int k = 0; k < 3; k++)
		{
			gc.position[k] = cam.position[k];
			gc.u[k] = u[k];
			gc.v[k] = v[k];
			gc.horizontal[k] = horizontal[k];
			gc.vertical[k] = vertical[k];
			gc.lowerLeftCorner[k] = llc[k];
		}
		return (gc);
	}

	namespace
	{
		// The atmosphere's default sun: LightUnits::solarDirectionalIrradiance of
		// ColorScience::solar() (Atmosphere.cpp:76, LightUnits.cpp:35-50 and 113-117), the
		// 5778 K blackbody colour scaled to a luminance of 1361 W/m^2.
		dvec3	solarSunRadiance()
		{
			dvec3 c = blackbodyColor(5778.0);
			const double scale = 1361.0 / luminance(c);
			for (double& v : c) v *= scale;
			return c;
		}

		GpuAtmosphere	packAtmosphere(const Scene& scene)
		{
			const Settings& s = scene.settings;
			GpuAtmosphere a{};
			// updateSunDirectionVector (Atmosphere.cpp:304-309), normalised (:32-44)
			const double angle = M_PI * s.sunAngle;
			double d[3] = {0.0, std::cos(angle), -std::sin(angle)};
			dvec3 sun = solarSunRadiance();
			// Scene::syncAtmosphereSunDirection (Scene.cpp:457-475): the first directional light
			// in hittable order is the sun, towards minus its direction, with its solar radiance
			// or else its emission
			for (const Hittable& h : scene.hittables)
			{
				if (h.kind != HIT_DIRECTIONAL) continue;
				const DirectionalLight& dl = scene.directionals[h.index];
				for (int k = 0; k < 3; k++) d[k] = dl.direction[k] * -1.0;
				sun = dl.hasSunRadiance ? dl.sunRadiance : scene.materials[h.material].color;
				break;
			}
			const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
			for (int k = 0; k < 3; k++)
			{
				a.sunDirection[k] = d[k] / len;
				a.sunRadiance[k] = sun[k] * s.sunRadianceScale;
			}
			a.earthRadius = s.earthRadius;
			a.atmosphereRadius = s.atmosphereRadius;
			a.inverseHR = 1.0 / s.atmosphereHR;
			a.inverseHM = 1.0 / s.atmosphereHM;
			a.starsBrightness = s.starsBrightness;
			a.metersPerUnit = s.metersPerUnit;
			a.samples = static_cast<std::uint32_t>(s.atmosphereSamples);
			a.lightSamples = static_cast<std::uint32_t>(s.atmosphereLightSamples);
			return a;
		}
	}

	PackedScene	packScene(const Scene& scene)
	{
		PackedScene p;
		const Settings& s = scene.settings;
		p.adaptive.enabled = s.adaptive;
		p.adaptive.minSamples = s.adaptiveMinSamples;
		p.adaptive.checkInterval = s.adaptiveCheckInterval;
		p.adaptive.threshold = s.adaptiveThreshold;
		p.post.viewTransform = s.viewTransform;
		p.post.bloom = s.bloom;
		p.post.exposure = s.exposure;
		p.post.contrast = s.contrast;

		// ---- materials -------------------------------------------------------------
		for (const Material& m : scene.materials)
		{
			GpuMaterial g{};
			g.kind = m.kind;
			for (int i = 0; i < 3; i++)
			{
				g.color[i] = static_cast<float>(m.color[i]);
				g.eta[i] = static_cast<float>(m.eta[i]);
				g.k[i] = static_cast<float>(m.k[i]);
				// a fully opaque interior (coefficient T_MAX) stays infinite in fp32
				g.absorption[i] = m.absorption[i] > static_cast<double>(FLT_MAX) ? INFINITY : static_cast<float>(m.absorption[i]);
			}
			// BSDF.cpp:54-58: alpha = max(1e-4, clamp01(r)^2)
			const double r = std::isfinite(m.roughness) ? std::max(0.0, std::min(1.0, m.roughness)) : 0.0;
			g.alpha = static_cast<float>(std::max(1e-4, r * r));
			g.roughness = static_cast<float>(m.roughness);
			g.ior = static_cast<float>(m.ior);
			g.anisotropy = static_cast<float>(m.anisotropy);
			auto alphaOf = [](double rough) {
				const double rr = std::isfinite(rough) ? std::max(0.0, std::min(1.0, rough)) : 0.0;
				return std::max(1e-4, rr * rr);
			};
			for (int i = 0; i < 3; i++) g.glossyColor[i] = static_cast<float>(m.glossyColor[i]);
			g.glossyWeight = static_cast<float>(m.glossyWeight);
			g.metallic = static_cast<float>(m.metallic);
			g.transmission = static_cast<float>(m.transmission);
			g.clearcoat = static_cast<float>(m.clearcoat);
			g.clearcoatAlpha = static_cast<float>(alphaOf(m.clearcoatRoughness));
			g.sheen = static_cast<float>(m.sheen);
			bool smoothTransmission = false;
			std::uint32_t subsurfaceFlags = 0u;
			if (m.kind == MAT_PRINCIPLED)
			{
				// componentWeights (Principled.cpp:158-202) for the material's constant base color
				auto c01 = [](double v) { return std::isfinite(v) ? std::max(0.0, std::min(1.0, v)) : 0.0; };
				const dvec3 base = {c01(m.color[0]), c01(m.color[1]), c01(m.color[2])};
				auto lumW = [](const dvec3& c) { const double l = luminance(c); return (!std::isfinite(l) || l <= 0.0) ? 0.0 : l; };
				const double r0 = (m.ior - 1.0) / (m.ior + 1.0);
				const double f0 = c01(r0 * r0);
				const double baseLum = std::max(0.05, lumW(base));
				const double diffuse = (1.0 - m.metallic) * (1.0 - m.transmission) * (1.0 - f0) * (1.0 - 0.25 * m.clearcoat);
				const dvec3 ssColor = {c01(m.subsurfaceColor[0]), c01(m.subsurfaceColor[1]), c01(m.subsurfaceColor[2])};
				const dvec3 ssAlbedo = {base[0] * ssColor[0], base[1] * ssColor[1], base[2] * ssColor[2]};
				const double w[7] = {
					diffuse * (1.0 - m.subsurface) * baseLum,
					diffuse * m.subsurface * std::max(0.05, lumW(ssAlbedo)),
					diffuse * m.sheen * 0.35,
					(1.0 - m.metallic) * std::max(0.03, f0) * (1.0 - 0.5 * m.transmission),
					m.metallic * std::max(0.05, baseLum),
					(1.0 - m.metallic) * m.transmission,
					m.clearcoat * 0.25 * (1.0 - 0.5 * m.roughness),
				};
				double total = 0.0;
				for (double x : w) if (std::isfinite(x) && x > 0.0) total += x;
				for (int i = 0; i < 7; i++) g.lobe[i] = total > 0.0 ? static_cast<float>(w[i] / total) : (i == 0 ? 1.0f : 0.0f);
				g.diffuseScale = static_cast<float>(diffuse);
				smoothTransmission = m.metallic <= 1e-6 && m.roughness <= 1e-4 && m.transmission >= 1.0 - 1e-6
					&& m.clearcoat <= 1e-6 && m.sheen <= 1e-6 && m.subsurface <= 1e-6;
				g.subsurface = static_cast<float>(m.subsurface);
				const dvec3 radius = m.subsurfaceRadius * m.subsurfaceScale;
				g.subsurfaceRadius = static_cast<float>(std::max(radius[0], std::max(radius[1], radius[2])));
				for (int i = 0; i < 3; i++) g.subsurfaceAlbedo[i] = static_cast<float>(ssAlbedo[i]);
				if (m.subsurface > 0.0)
				{
					subsurfaceFlags = m.subsurfaceThin ? GPU_MATF_THIN_SUBSURFACE : GPU_MATF_BURLEY_SUBSURFACE;
				}
			}
			p.materialColor.push_back(m.color);
			g.flags = subsurfaceFlags | (smoothTransmission ? GPU_MATF_SMOOTH_TRANSMISSION : 0u) | (m.conductor ? GPU_MATF_CONDUCTOR : 0u)
				| ((m.absorption[0] > 0.0 || m.absorption[1] > 0.0 || m.absorption[2] > 0.0) ? GPU_MATF_ABSORPTION : 0u);
			p.materials.push_back(g);
		}

		// ---- meshes ------------------------------------------------------------------
		std::vector<std::int32_t> meshRoot(scene.meshes.size()), meshTriBase(scene.meshes.size()), meshNormalBase(scene.meshes.size(), -1);
		p.blas.reserve(scene.meshes.size());
		for (std::size_t m = 0; m < scene.meshes.size(); m++)
		{
			const Mesh& mesh = scene.meshes[m];
			p.blas.push_back(buildBlas(mesh));
			const Blas& blas = p.blas.back();
			const std::int32_t triBase = static_cast<std::int32_t>(p.meshTris.size());
			meshTriBase[m] = triBase;
			bool anyNormals = false;
			for (const Triangle& t : mesh.triangles)
			{
				anyNormals = anyNormals || t.hasNormals;
			}
			if (anyNormals)
			{
				meshNormalBase[m] = static_cast<std::int32_t>(p.meshNormals.size() / 9);
			}
			for (std::size_t k = 0; k < blas.triangleIndices.size(); k++)
			{
				const Triangle& t = mesh.triangles[blas.triangleIndices[k]];
				p.meshTris.push_back(packTriangle(t, static_cast<std::int32_t>(blas.triangleIndices[k])));
				if (anyNormals)
				{
					const dvec3* ns[3] = {&t.n0, &t.n1, &t.n2};
					for (int v = 0; v < 3; v++)
						for (int i = 0; i < 3; i++)
							p.meshNormals.push_back(t.hasNormals ? static_cast<float>((*ns[v])[i]) : 0.0f);
				}
			}
			const std::int32_t binaryRoot = packBlas(blas, triBase, p.blasNodes);
			meshRoot[m] = widen(p.blasNodes, binaryRoot, p.blasWide);
			p.blasDeviceDepth = std::max(p.blasDeviceDepth, wideDepth(p.blasWide, meshRoot[m]));
			p.blasStackBound = std::max(p.blasStackBound, wideStackBound(p.blasWide, meshRoot[m]));
			p.triangleCount += mesh.triangles.size();
		}

		// ---- top-level hittables -----------------------------------------------------
		const std::size_t nh = scene.hittables.size();
		std::vector<AABBd> boxes(nh);
		std::vector<bool> bounded(nh, false);
		for (std::size_t i = 0; i < nh; i++)
		{
			const Hittable& h = scene.hittables[i];
			if (h.material >= scene.materials.size())
			{
				throw std::runtime_error("Hittable references an unknown material");
			}
			GpuHittable gh{};
			gh.material = h.material;
			switch (h.kind)
			{
				case HIT_TRIANGLE:
				{
					const Triangle& t = scene.triangles[h.index];
					boxes[i] = triangleBox(t);
					bounded[i] = true;
					gh.kindIndex = (GPU_HIT_TRIANGLE << kKindShift) | static_cast<std::uint32_t>(p.topTris.size());
					p.topTris.push_back(packTriangle(t, static_cast<std::int32_t>(i)));
					p.triangleCount++;
					break;
				}
				case HIT_RECTANGLE:
				{
					const Rectangle& r = scene.rectangles[h.index];
					bounded[i] = rectangleBox(r, boxes[i]);
					gh.kindIndex = (GPU_HIT_RECTANGLE << kKindShift) | static_cast<std::uint32_t>(p.rects.size());
					p.rects.push_back(packRect(r));
					break;
				}
				case HIT_SPHERE:
				{
					const Sphere& sp = scene.spheres[h.index];
					boxes[i] = sphereBox(sp.position, sp.radius);
					bounded[i] = true;
					GpuSphere g{};
					for (int k = 0; k < 3; k++) g.center[k] = static_cast<float>(sp.position[k]);
					g.radius = static_cast<float>(sp.radius);
					gh.kindIndex = (GPU_HIT_SPHERE << kKindShift) | static_cast<std::uint32_t>(p.spheres.size());
					p.spheres.push_back(g);
					break;
				}
				case HIT_CUBE:
				{
					const Cube& c = scene.cubes[h.index];
					boxes[i] = boxAround(c.position, c.width, c.height, c.depth);
					bounded[i] = true;
					gh.kindIndex = (GPU_HIT_CUBE << kKindShift) | static_cast<std::uint32_t>(p.cubeFaces.size() / 6);
					for (const Rectangle& f : cubeFaces(c)) p.cubeFaces.push_back(packRect(f));
					break;
				}
				case HIT_VOLUME:
				{
					const Volume& v = scene.volumes[h.index];
					boxes[i] = v.sphere ? sphereBox(v.position, v.radius) : boxAround(v.position, v.size[0], v.size[1], v.size[2]);
					bounded[i] = true;
					GpuVolume g{};
					for (int k = 0; k < 3; k++)
					{
						g.center[k] = static_cast<float>(v.position[k]);
						g.lo[k] = static_cast<float>(boxes[i].lo[k]);
						g.hi[k] = static_cast<float>(boxes[i].hi[k]);
					}
					g.radius = static_cast<float>(v.radius);
					g.sphere = v.sphere ? 1u : 0u;
					g.negInvDensity = static_cast<float>(-1.0 / v.density);
					gh.kindIndex = (GPU_HIT_VOLUME << kKindShift) | static_cast<std::uint32_t>(p.volumes.size());
					p.volumes.push_back(g);
					break;
				}
				case HIT_PLANE:
				{
					const Plane& pl = scene.planes[h.index];
					GpuPlane g{};
					g.point[0] = 0.0f;
					g.point[1] = static_cast<float>(pl.y);
					g.point[2] = 0.0f;
					for (int k = 0; k < 3; k++) g.orientation[k] = static_cast<float>(pl.orientation[k]);
					gh.kindIndex = (GPU_HIT_PLANE << kKindShift) | static_cast<std::uint32_t>(p.planes.size());
					p.planes.push_back(g);
					break;
				}
				case HIT_DIRECTIONAL:
				{
					gh.kindIndex = (GPU_HIT_DIRECTIONAL << kKindShift) | h.index;
					break;
				}
				case HIT_INSTANCE:
				{
					const Instance& inst = scene.instances[h.index];
					bounded[i] = instanceBox(inst, p.blas[inst.mesh], boxes[i]);
					const InstanceXf xf(inst);
					// world -> local: inverseRotate(p - position) / scale as a 3x4 matrix
					auto inverseRotate = [&](dvec3 v) {
						// MeshInstance.cpp:100-121
						double x = v[0], y = v[1], z = v[2];
						double nx = (x * xf.cz) + (y * xf.sz);
						double ny = (-x * xf.sz) + (y * xf.cz);
						x = nx; y = ny;
						nx = (x * xf.cy) - (z * xf.sy);
						double nz = (x * xf.sy) + (z * xf.cy);
						x = nx; z = nz;
						ny = (y * xf.cx) + (z * xf.sx);
						nz = (-y * xf.sx) + (z * xf.cx);
						return dvec3{x, ny, nz};
					};
					GpuInstance g{};
					double cols[3][3];
					for (int c = 0; c < 3; c++)
					{
						dvec3 e{}; e[c] = 1.0;
						const dvec3 r = inverseRotate(e);
						for (int k = 0; k < 3; k++) cols[c][k] = r[k] / inst.scale[k];
					}
					const dvec3 pr = inverseRotate(inst.position);
					for (int row = 0; row < 3; row++)
					{
						for (int c = 0; c < 3; c++) g.worldToLocal[row * 4 + c] = static_cast<float>(cols[c][row]);
						g.worldToLocal[row * 4 + 3] = static_cast<float>(-pr[row] / inst.scale[row]);
					}
					// normal to world: rotate(n / scale)
					for (int c = 0; c < 3; c++)
					{
						dvec3 e{}; e[c] = 1.0 / inst.scale[c];
						const dvec3 r = xf.rotate(e);
						for (int row = 0; row < 3; row++) g.normalToWorld[row * 3 + c] = static_cast<float>(r[row]);
					}
					g.blasRoot = meshRoot[inst.mesh];
					g.triBase = meshTriBase[inst.mesh];
					g.normalBase = meshNormalBase[inst.mesh];
					gh.kindIndex = (GPU_HIT_INSTANCE << kKindShift) | static_cast<std::uint32_t>(p.instances.size());
					p.instances.push_back(g);
					break;
				}
			}
			p.hittables.push_back(gh);
		}

		// ---- top-level BVH (Scene.cpp:655-688) ------------------------------------------
		std::vector<std::uint32_t> items;
		for (std::size_t i = 0; i < nh; i++)
		{
			if (bounded[i]) items.push_back(static_cast<std::uint32_t>(i));
			else p.unbounded.push_back(static_cast<std::uint32_t>(i));
		}
		// the device tests planes after the BVH (HitHelper.cpp:5-32); a directional light is
		// never hit, and a rectangle without a box is degenerate and is never hit either
		// (Rectangle::hit rejects it through the same basis test)
		{
			std::vector<std::uint32_t> tested;
			for (std::uint32_t u : p.unbounded)
			{
				if ((p.hittables[u].kindIndex >> kKindShift) == GPU_HIT_PLANE) tested.push_back(u);
			}
			p.unbounded.swap(tested);
		}
		p.params.tlasRoot = -1;
		if (!items.empty())
		{
			TlasBuilder tb{boxes, {}};
			if (items.size() == 1)
			{
				// Scene.cpp:681-684: a single bounded hittable is the acceleration structure
				TopNode leaf;
				leaf.children = {items[0]};
				leaf.childBoxes = {boxes[items[0]]};
				leaf.box = boxes[items[0]];
				tb.nodes.push_back(leaf);
				p.tlasTopology = std::to_string(items[0]);
			}
			else
			{
				tb.build(items, 0, items.size(), true);
				std::ostringstream os;
				tb.dump(os, 0);
				p.tlasTopology = os.str();
			}
			// Device form: the reference's binary upper levels are kept node for node; each
			// n-ary leaf (up to 16 hittables) is refined into a small binary tree so every
			// mesh instance has its own box test before its mesh BVH is entered, and plain
			// primitives sit in leaves of at most 4.
			TlasEmitter em{boxes, p.hittables, p.tlasNodes, p.tlasPrims};
			const TlasEmitter::Ref root = em.fromReference(tb, 0);
			if (root.count > 0)
			{
				BvhNode r{};
				setChild(r, 0, &root.box, root.child, root.count);
				setChild(r, 1, nullptr, 0, -1);
				p.tlasNodes.push_back(r);
				p.params.tlasRoot = static_cast<std::int32_t>(p.tlasNodes.size() - 1);
			}
			else
			{
				p.params.tlasRoot = root.child;
			}
			p.params.tlasRoot = widen(p.tlasNodes, p.params.tlasRoot, p.tlasWide);
			p.tlasDeviceDepth = wideDepth(p.tlasWide, p.params.tlasRoot);
			p.tlasStackBound = wideStackBound(p.tlasWide, p.params.tlasRoot);
		}
		p.tlasLeaf.reserve(2 * p.tlasPrims.size());
		for (const std::uint32_t h : p.tlasPrims)
		{
			p.tlasLeaf.push_back(h);
			std::uint32_t kindIndex = p.hittables[h].kindIndex;
			if (scene.hittables[h].kind == HIT_SPHERE && !scene.spheres[scene.hittables[h].index].visible)
			{
				kindIndex = (GPU_HIT_HIDDEN_SPHERE << kKindShift) | (kindIndex & kIndexMask);
			}
			p.tlasLeaf.push_back(kindIndex);
		}

		// ---- lights (Scene.cpp:615-638) -----------------------------------------------------
		double total = 0.0;
		std::vector<double> weights;
		for (std::size_t i = 0; i < nh; i++)
		{
			const Hittable& h = scene.hittables[i];
			const Material& m = scene.materials[h.material];
			if (m.kind != MAT_EMISSIVE)
			{
				continue;
			}
			GpuLight l{};
			l.hittable = static_cast<std::uint32_t>(i);
			l.material = h.material;
			for (int k = 0; k < 3; k++) l.emitted[k] = static_cast<float>(m.color[k]);
			const double lum = luminance(m.color);
			double area = 0.0;
			if (h.kind == HIT_TRIANGLE)
			{
				const Triangle& t = scene.triangles[h.index];
				const dvec3 c = cross(t.v1 - t.v0, t.v2 - t.v0);
				const double ls = dot(c, c);
				area = ls <= 1e-24 ? 0.0 : std::sqrt(ls) / 2.0;
				l.kind = GPU_HIT_TRIANGLE;
				l.index = p.hittables[i].kindIndex & kIndexMask;
			}
			else if (h.kind == HIT_RECTANGLE)
			{
				const Rectangle& r = scene.rectangles[h.index];
				area = r.width * r.height;
				l.kind = GPU_HIT_RECTANGLE;
				l.index = p.hittables[i].kindIndex & kIndexMask;
			}
			else if (h.kind == HIT_SPHERE)
			{
				// Sphere::lightSelectionWeight (Sphere.cpp:501-515)
				const Sphere& sp = scene.spheres[h.index];
				area = 4.0 * kPi * sp.radius * sp.radius;
				l.kind = GPU_HIT_SPHERE;
				l.index = p.hittables[i].kindIndex & kIndexMask;
				if (sp.ies >= 0)
				{
					// relativeIntensity's axis (IESProfile.cpp:302-305): normalised, or straight
					// down when it has no length
					const IesProfile& pr = scene.iesProfiles[static_cast<std::size_t>(sp.ies)];
					dvec3 axis = sp.iesDirection;
					const double l2 = dot(axis, axis);
					if (!std::isfinite(l2) || l2 <= 0.0) axis = dvec3{0.0, -1.0, 0.0};
					else
					{
						const double len = std::sqrt(l2);
						axis = dvec3{axis[0] / len, axis[1] / len, axis[2] / len};
					}
					l.ies = static_cast<std::uint32_t>(p.iesData.size()) + 1u;
					p.iesData.insert(p.iesData.end(), {pr.meanCandela, axis[0], axis[1], axis[2], sp.iesRotation,
						static_cast<double>(pr.vertical.size()), static_cast<double>(pr.horizontal.size())});
					p.iesData.insert(p.iesData.end(), pr.vertical.begin(), pr.vertical.end());
					p.iesData.insert(p.iesData.end(), pr.horizontal.begin(), pr.horizontal.end());
					p.iesData.insert(p.iesData.end(), pr.candela.begin(), pr.candela.end());
				}
			}
			else if (h.kind == HIT_DIRECTIONAL)
			{
				// DirectionalLight::sampleLight / lightSelectionWeight (DirectionalLight.cpp:74-103)
				const DirectionalLight& d = scene.directionals[h.index];
				for (int k = 0; k < 3; k++) l.direction[k] = static_cast<float>(d.direction[k] * -1.0);
				l.kind = GPU_HIT_DIRECTIONAL;
				area = 1.0;
			}
			else if (h.kind == HIT_CUBE || h.kind == HIT_PLANE)
			{
				// Hittable::sampleLight / pdfValue defaults (Hittable.cpp): never a valid
				// sample, weight = luminance
				l.kind = h.kind == HIT_CUBE ? GPU_HIT_CUBE : GPU_HIT_PLANE;
				area = 1.0;
			}
			else if (h.kind == HIT_INSTANCE)
			{
				// MeshInstance::_surfaceArea / _transformedAreaPrefixSums / lightSelectionWeight
				// (MeshInstance.cpp:345-407,589-603): a uniform scale keeps the mesh's own
				// prefix sums (Mesh.cpp:525-532) and scales the total, any other scale sums the
				// transformed triangle areas
				const Instance& inst = scene.instances[h.index];
				const Mesh& mesh = scene.meshes[inst.mesh];
				const InstanceXf xf(inst);
				const double ax = std::fabs(inst.scale[0]), ay = std::fabs(inst.scale[1]), az = std::fabs(inst.scale[2]);
				const bool uniform = std::fabs(ax - ay) <= 1e-12 && std::fabs(ax - az) <= 1e-12;
				const std::vector<std::size_t>& order = p.blas[inst.mesh].triangleIndices;
				std::vector<std::uint32_t> slotOf(order.size());
				for (std::size_t j = 0; j < order.size(); j++) slotOf[order[j]] = static_cast<std::uint32_t>(meshTriBase[inst.mesh] + j);
				l.kind = GPU_HIT_INSTANCE;
				l.index = p.hittables[i].kindIndex & kIndexMask;
				l.triBase = static_cast<std::uint32_t>(p.lightTriCdf.size());
				l.triCount = static_cast<std::uint32_t>(mesh.triangles.size());
				double total = 0.0;
				for (std::size_t k = 0; k < mesh.triangles.size(); k++)
				{
					const Triangle& t = mesh.triangles[k];
					double a;
					if (uniform)
					{
						const dvec3 c = cross(t.v1 - t.v0, t.v2 - t.v0);
						const double ls = dot(c, c);
						a = ls <= 1e-24 ? 0.0 : std::sqrt(ls) / 2.0;	// Triangle::_updateCache
					}
					else
					{
						const dvec3 w0 = xf.transformPoint(t.v0), w1 = xf.transformPoint(t.v1), w2 = xf.transformPoint(t.v2);
						const dvec3 c = cross(w1 - w0, w2 - w0);
						a = std::sqrt(dot(c, c)) / 2.0;
						if (!std::isfinite(a) || a <= 0.0) a = 0.0;
					}
					total += a;
					p.lightTriCdf.push_back(total);
					p.lightTriSlot.push_back(slotOf[k]);
				}
				if (!std::isfinite(total) || total <= 0.0) area = 0.0;
				else area = uniform ? total * (ax * ax) : total;
				if (!std::isfinite(area) || area <= 0.0) area = 0.0;
				// local -> world: rotate(p * scale) + position, as a 3x4 map
				for (int c = 0; c < 3; c++)
				{
					dvec3 e{};
					e[c] = inst.scale[c];
					const dvec3 r = xf.rotate(e);
					for (int row = 0; row < 3; row++) l.localToWorld[row * 4 + c] = static_cast<float>(r[row]);
				}
				for (int row = 0; row < 3; row++) l.localToWorld[row * 4 + 3] = static_cast<float>(inst.position[row]);
			}
			else
			{
				throw std::runtime_error("Unsupported for the device path: emissive volume");
			}
			l.area = static_cast<float>(h.kind == HIT_DIRECTIONAL || h.kind == HIT_CUBE || h.kind == HIT_PLANE ? 0.0 : area);
			p.lightScale.push_back(area);
			double w = (area <= 0.0 || !std::isfinite(lum) || lum <= 0.0) ? 0.0 : area * lum;
			if (!std::isfinite(w) || w <= 0.0) w = 0.0;
			total += w;
			weights.push_back(w);
			p.lightCdf.push_back(total);
			p.lights.push_back(l);
		}
		if (total <= 0.0)
		{
			p.lightCdf.clear();
		}
		for (std::size_t i = 0; i < p.lights.size(); i++)
		{
			// lightSelectionProbability (ColorHelper.cpp:176-199)
			double prob;
			if (p.lightCdf.empty())
			{
				prob = 1.0 / static_cast<double>(p.lights.size());
			}
			else
			{
				const double prev = i == 0 ? 0.0 : p.lightCdf[i - 1];
				const double w = p.lightCdf[i] - prev;
				prob = (w <= 0.0 || total <= 0.0) ? 0.0 : w / total;
			}
			p.lights[i].selectionProbability = static_cast<float>(prob);
		}

		// ---- environment (EnvironmentMap.cpp:627-671) ----------------------------------------
		GpuParams& gp = p.params;
		gp.hasEnvironment = scene.hasEnvironment ? 1u : 0u;
		gp.envTotalWeight = 0.0;
		if (scene.hasEnvironment)
		{
			const Environment& env = scene.environment;
			const std::size_t W = env.width, H = env.height;
			p.envPixels.resize(W * H * 4);
			p.envPdf.assign(W * H, 0.0f);
			p.envCdf.reserve(W * H);
			std::vector<double> weightsE(W * H), solid(W * H);
			double tw = 0.0;
			const double deltaPhi = 2.0 * kPi / static_cast<double>(W);
			for (std::size_t y = 0; y < H; y++)
			{
				const double t0 = kPi * static_cast<double>(y) / static_cast<double>(H);
				const double t1 = kPi * static_cast<double>(y + 1) / static_cast<double>(H);
				const double sa = deltaPhi * std::max(0.0, std::cos(t0) - std::cos(t1));
				for (std::size_t x = 0; x < W; x++)
				{
					const std::size_t i = y * W + x;
					const dvec3& c = env.pixels[i];
					dvec3 sc;
					for (int k = 0; k < 3; k++) sc[k] = std::isfinite(c[k]) ? std::max(0.0, c[k]) : 0.0;
					const double lum = luminance(sc);
					const double w = (std::isfinite(lum) && lum > 0.0) ? lum * sa : 0.0;
					solid[i] = sa;
					weightsE[i] = w;
					tw += w;
					p.envCdf.push_back(tw);
					for (int k = 0; k < 3; k++) p.envPixels[i * 4 + k] = static_cast<float>(c[k]);
					p.envPixels[i * 4 + 3] = 0.0f;
				}
			}
			if (!std::isfinite(tw) || tw <= 0.0)
			{
				tw = 0.0;
				p.envCdf.clear();
			}
			else
			{
				for (std::size_t i = 0; i < W * H; i++)
				{
					if (weightsE[i] > 0.0 && solid[i] > 0.0)
					{
						p.envPdf[i] = static_cast<float>((weightsE[i] / tw) / solid[i]);
					}
				}
				const std::uint32_t n = static_cast<std::uint32_t>(W * H);
				if (n >= kEnvGuideMinTexels)
				{
					// the bucket bounds are formed as the sampler forms its target:
					// (g / G exactly) x total, one rounding
					const std::uint32_t G = envGuideBuckets(n);
					p.envGuide.assign(n, 0u);
					for (std::uint32_t g = 0; g <= G; g++)
					{
						const double target = (static_cast<double>(g) / static_cast<double>(G)) * tw;
						p.envGuide[g] = static_cast<std::uint32_t>(std::lower_bound(p.envCdf.begin(), p.envCdf.end(), target) - p.envCdf.begin());
					}
				}
			}
			gp.envTotalWeight = tw;
			gp.envWidth = static_cast<std::uint32_t>(W);
			gp.envHeight = static_cast<std::uint32_t>(H);
			gp.envStrength = static_cast<float>(env.strength);
			gp.envRotation = static_cast<float>(env.rotation);
			gp.envLighting = env.lighting ? 1u : 0u;
		}

		gp.camera = makeCamera(scene.camera, s.width, s.height, s.metersPerUnit);
		p.camera = scene.camera;
		p.metersPerUnit = s.metersPerUnit;
		gp.width = s.width;
		gp.height = s.height;
		gp.maxBounces = s.maxBounces;
		gp.totalSamples = s.samples;
		gp.sky = s.sky == SKY_ENVIRONMENT ? GPU_SKY_ENVIRONMENT : s.sky == SKY_LINEAR ? GPU_SKY_LINEAR
			: s.sky == SKY_ATMOSPHERE ? GPU_SKY_ATMOSPHERE : GPU_SKY_NONE;
		if (s.sky == SKY_ATMOSPHERE) gp.atmosphere = packAtmosphere(scene);
		for (int k = 0; k < 3; k++) gp.background[k] = static_cast<float>(s.background[k]);
		gp.backgroundLuminance = static_cast<float>(luminance(s.background));
		gp.lightCount = static_cast<std::uint32_t>(p.lights.size());
		gp.metersPerUnit = static_cast<float>(s.metersPerUnit);
		gp.unboundedCount = static_cast<std::uint32_t>(p.unbounded.size());
		gp.hasVolumes = p.volumes.empty() ? 0u : 1u;
		// a scene with denoising on samples on streams 1 / 2 (Threads.cpp:594-598), also when
		// only its raw image is rendered
		gp.denoiseStreams = s.denoise ? 1u : 0u;
		p.features = 0;
		for (const Hittable& h : scene.hittables)
		{
			if (h.kind == HIT_SPHERE || h.kind == HIT_CUBE || h.kind == HIT_VOLUME || h.kind == HIT_PLANE || h.kind == HIT_DIRECTIONAL
				|| (h.kind == HIT_INSTANCE && scene.materials[h.material].kind == MAT_EMISSIVE))
			{
				p.features |= kFeatAnalytic;	// the class of rarer primitives and lights
			}
		}
		for (const Material& m : scene.materials)
		{
			if (m.kind == MAT_GLOSSY || m.kind == MAT_DIFFUSE_GLOSSY || m.kind == MAT_PRINCIPLED || m.kind == MAT_ISOTROPIC || m.kind == MAT_HG)
			{
				p.features |= kFeatRich;
			}
			if (m.kind == MAT_PRINCIPLED && m.subsurface > 0.0 && !m.subsurfaceThin)
			{
				p.features |= kFeatEvery;	// the relocation probe is compiled only into these kernels
			}
		}
		if (s.sky == SKY_ATMOSPHERE) p.features |= kFeatRich;	// the atmosphere march lives in the rich kernels
		gp.lightWeighted = p.lightCdf.empty() ? 0u : 1u;
		gp.lightBase = 0;
		gp.envBase = 0;
		gp.unboundedBase = 0;
		if (p.unbounded.size() > 0xffffu)
		{
			throw std::runtime_error("Unsupported for the device path: more than 65535 planes");
		}
		return (p);
	}
}

namespace luzgpu
{
	// Concatenates packed scenes into one set of device arrays (a multi-scene launch), and
	// returns each scene's parameters with its TLAS root and table bases in the merged arrays.
	PackedScene	mergeScenes(const std::vector<const PackedScene*>& parts, std::vector<GpuParams>& params)
	{
		PackedScene m;
		params.clear();
		for (const PackedScene* part : parts)
		{
			const PackedScene& p = *part;
			const std::uint32_t tlasBase = static_cast<std::uint32_t>(m.tlasWide.size());
			const std::uint32_t blasBase = static_cast<std::uint32_t>(m.blasWide.size());
			const std::uint32_t meshTriBase = static_cast<std::uint32_t>(m.meshTris.size());
			const std::uint32_t normalSlotBase = static_cast<std::uint32_t>(m.meshNormals.size() / 9);
			const std::uint32_t leafBase = static_cast<std::uint32_t>(m.tlasLeaf.size() / 2);
			const std::uint32_t hittableBase = static_cast<std::uint32_t>(m.hittables.size());
			const std::uint32_t materialBase = static_cast<std::uint32_t>(m.materials.size());
			const std::uint32_t base[9] = {
				static_cast<std::uint32_t>(m.topTris.size()), static_cast<std::uint32_t>(m.rects.size()),
				static_cast<std::uint32_t>(m.instances.size()), static_cast<std::uint32_t>(m.spheres.size()),
				static_cast<std::uint32_t>(m.cubeFaces.size() / 6), static_cast<std::uint32_t>(m.volumes.size()),
				static_cast<std::uint32_t>(m.planes.size()), 0u, static_cast<std::uint32_t>(m.spheres.size())};
			auto rebaseKind = [&](std::uint32_t kindIndex) {
				const std::uint32_t kind = kindIndex >> kKindShift;
				const std::uint32_t idx = (kindIndex & kIndexMask) + base[kind];
				if (idx > kIndexMask) throw std::runtime_error("Unsupported: batch too large for 28-bit primitive indices");
				return (kind << kKindShift) | idx;
			};
			for (Bvh4Node n : p.tlasWide)
			{
				for (int k = 0; k < 4; k++)
				{
					if (n.count[k] == 0) n.child[k] += static_cast<std::int32_t>(tlasBase);
					else if (n.count[k] != 255) n.child[k] += static_cast<std::int32_t>(leafBase);
				}
				m.tlasWide.push_back(n);
			}
			for (Bvh4Node n : p.blasWide)
			{
				for (int k = 0; k < 4; k++)
				{
					if (n.count[k] == 0) n.child[k] += static_cast<std::int32_t>(blasBase);
					else if (n.count[k] != 255) n.child[k] += static_cast<std::int32_t>(meshTriBase);
				}
				m.blasWide.push_back(n);
			}
			m.meshTris.insert(m.meshTris.end(), p.meshTris.begin(), p.meshTris.end());
			m.meshNormals.insert(m.meshNormals.end(), p.meshNormals.begin(), p.meshNormals.end());
			for (std::size_t i = 0; i + 1 < p.tlasLeaf.size(); i += 2)
			{
				m.tlasLeaf.push_back(p.tlasLeaf[i] + hittableBase);
				m.tlasLeaf.push_back(rebaseKind(p.tlasLeaf[i + 1]));
			}
			for (GpuHittable h : p.hittables)
			{
				h.kindIndex = rebaseKind(h.kindIndex);
				h.material += materialBase;
				m.hittables.push_back(h);
			}
			for (GpuTri t : p.topTris)
			{
				t.prim += static_cast<std::int32_t>(hittableBase);
				m.topTris.push_back(t);
			}
			m.rects.insert(m.rects.end(), p.rects.begin(), p.rects.end());
			for (GpuInstance in : p.instances)
			{
				in.blasRoot += static_cast<std::int32_t>(blasBase);
				in.triBase += static_cast<std::int32_t>(meshTriBase);
				if (in.normalBase >= 0) in.normalBase += static_cast<std::int32_t>(normalSlotBase);
				m.instances.push_back(in);
			}
			m.spheres.insert(m.spheres.end(), p.spheres.begin(), p.spheres.end());
			m.cubeFaces.insert(m.cubeFaces.end(), p.cubeFaces.begin(), p.cubeFaces.end());
			m.volumes.insert(m.volumes.end(), p.volumes.begin(), p.volumes.end());
			m.planes.insert(m.planes.end(), p.planes.begin(), p.planes.end());
			m.materials.insert(m.materials.end(), p.materials.begin(), p.materials.end());
			GpuParams gp = p.params;
			gp.lightBase = static_cast<std::uint32_t>(m.lights.size());
			const std::uint32_t lightTriBase = static_cast<std::uint32_t>(m.lightTriCdf.size());
			const std::uint32_t iesBase = static_cast<std::uint32_t>(m.iesData.size());
			m.iesData.insert(m.iesData.end(), p.iesData.begin(), p.iesData.end());
			for (GpuLight l : p.lights)
			{
				if (l.kind == GPU_HIT_TRIANGLE || l.kind == GPU_HIT_RECTANGLE || l.kind == GPU_HIT_SPHERE || l.kind == GPU_HIT_INSTANCE)
				{
					l.index += base[l.kind];
				}
				if (l.kind == GPU_HIT_INSTANCE) l.triBase += lightTriBase;
				if (l.ies) l.ies += iesBase;
				l.hittable += hittableBase;
				l.material += materialBase;
				m.lights.push_back(l);
			}
			m.lightTriCdf.insert(m.lightTriCdf.end(), p.lightTriCdf.begin(), p.lightTriCdf.end());
			for (std::uint32_t slot : p.lightTriSlot) m.lightTriSlot.push_back(slot + meshTriBase);
			// lightCdf is indexed like the lights: pad an unweighted scene's span
			if (p.lightCdf.empty()) m.lightCdf.insert(m.lightCdf.end(), p.lights.size(), 0.0);
			else m.lightCdf.insert(m.lightCdf.end(), p.lightCdf.begin(), p.lightCdf.end());
			gp.envBase = static_cast<std::uint32_t>(m.envPdf.size());
			m.envPixels.insert(m.envPixels.end(), p.envPixels.begin(), p.envPixels.end());
			m.envPdf.insert(m.envPdf.end(), p.envPdf.begin(), p.envPdf.end());
			// envCdf is empty when the map has no weight: keep it texel-aligned with envPdf
			if (p.envCdf.empty()) m.envCdf.insert(m.envCdf.end(), p.envPdf.size(), 0.0);
			else m.envCdf.insert(m.envCdf.end(), p.envCdf.begin(), p.envCdf.end());
			if (p.envGuide.empty()) m.envGuide.insert(m.envGuide.end(), p.envPdf.size(), 0u);
			else m.envGuide.insert(m.envGuide.end(), p.envGuide.begin(), p.envGuide.end());
			gp.unboundedBase = static_cast<std::uint32_t>(m.unbounded.size());
			for (std::uint32_t u : p.unbounded) m.unbounded.push_back(u + hittableBase);
			if (m.unbounded.size() > 0xffffu) throw std::runtime_error("Unsupported: more than 65535 planes in a batch");
			gp.tlasRoot = p.params.tlasRoot >= 0 ? p.params.tlasRoot + static_cast<std::int32_t>(tlasBase) : -1;
			params.push_back(gp);
			m.features |= p.features;
			m.tlasStackBound = std::max(m.tlasStackBound, p.tlasStackBound);
			m.blasStackBound = std::max(m.blasStackBound, p.blasStackBound);
			m.triangleCount += p.triangleCount;
		}
		return (m);
	}
}

namespace luzgpu
{
	void	rebuildLights(PackedScene& p)
	{
		double total = 0.0;
		p.lightCdf.clear();
		for (std::size_t i = 0; i < p.lights.size(); i++)
		{
			GpuLight& l = p.lights[i];
			const dvec3& c = p.materialColor[l.material];
			for (int k = 0; k < 3; k++) l.emitted[k] = static_cast<float>(c[k]);
			const double lum = luminance(c);
			const double area = p.lightScale[i];
			double w = (area <= 0.0 || !std::isfinite(lum) || lum <= 0.0) ? 0.0 : area * lum;
			if (!std::isfinite(w) || w <= 0.0) w = 0.0;
			total += w;
			p.lightCdf.push_back(total);
		}
		if (total <= 0.0) p.lightCdf.clear();
		for (std::size_t i = 0; i < p.lights.size(); i++)
		{
			double prob;
			if (p.lightCdf.empty()) prob = 1.0 / static_cast<double>(p.lights.size());
			else
			{
				const double prev = i == 0 ? 0.0 : p.lightCdf[i - 1];
				const double w = p.lightCdf[i] - prev;
				prob = (w <= 0.0 || total <= 0.0) ? 0.0 : w / total;
			}
			p.lights[i].selectionProbability = static_cast<float>(prob);
		}
		p.params.lightWeighted = p.lightCdf.empty() ? 0u : 1u;
	}
}
