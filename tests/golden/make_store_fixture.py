"""Writes tests/golden/store/: two graph directories in the reference's
storage format, with every array worked out by hand from the importer's
steps (no library or oracle code runs here).

knows/ -- LDBC-SNB-like undirected Person-knows-Person, V = 5 (vertex 4
isolated), raw edges in file order
    (0,1) (0,1) (1,2) (2,0) (0,1) (3,3)
  add_reverse_edges appends (1,0) (1,0) (2,1) (0,2) (1,0) (3,3)
  (graph_concepts.hh:96-103); deduplicate_edges is std::unique over that
  unsorted list (graph_concepts.hh:104-107), dropping only the 2nd (0,1) and
  the 2nd (1,0):
    (0,1) (1,2) (2,0) (0,1) (3,3) (1,0) (2,1) (0,2) (1,0) (3,3)   count 10
  sorted by (src, dst):
    (0,1) (0,1) (0,2) (1,0) (1,0) (1,2) (2,0) (2,1) (3,3) (3,3)
  CSR index [0,3,6,8,10,10], data [1,1,2,0,0,2,0,1,3,3]
  (sort_and_generate_index_from_src, graph_concepts.hh:114-137).
  Files: "Person-knows-Person_csr_src_dst[_index]", the flat
  "Person-knows-Person_src_dst" (VIDPair, the sorted pairs) and "Person.id".

fin/ -- LDBC-FinBench-like directed Account-transfer->Account with
multi-edges allowed (no dedup), V = 3, raw edges
    (0,1) (0,1) (2,0) (1,2) (0,1)
  forward, sorted by (src,dst): (0,1)x3 (1,2) (2,0)
    index [0,3,4,5], data [1,1,1,2,0]
  backward, sorted by (dst,src): (2,0) (0,1)x3 (1,2)
    index [0,1,4,5], data (source ids) [2,0,0,0,1]
  plus a Medium-signIn->Account relation (2 Media) between two vertex types:
  edges (0,2) (1,2) (1,0) -> forward index [0,1,3] data [2,0,2];
  backward (by Account) index [0,1,1,3] data [1,0,1].

Every DiskArray file is 64 header bytes {int64 count, int64 element size,
zeros} + the elements + zeros up to 64 + capacity * size, capacity
(4096 - 64) / size for these small arrays (disk_array.hh:267-354, 537-604),
so each file is 4096 bytes.  schema.json follows the field names of the
NLOHMANN_DEFINE_TYPE_INTRUSIVE lists (schema.hh:32, 121-123, 137-138) with
nlohmann's sorted keys and 4-space indent; FlatEdgeOrder has no free to_json,
so store_info holds its integer values (SrcDst = 1).
"""
import json
import os
import struct

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "store")


def diskarray(path, fmt, values):
    size = struct.calcsize("<" + fmt)
    body = b"".join(struct.pack("<" + fmt, *(v if isinstance(v, tuple) else (v,)))
                    for v in values)
    head = struct.pack("<qq", len(values), size) + b"\0" * 48
    with open(path, "wb") as f:
        f.write(head + body + b"\0" * (4096 - 64 - len(body)))


def vertex(type_, count):
    return {"properties": {"id": {"name": "id", "type": "VID"}},
            "statistics": {"count": count}, "type": type_}


def edge(type_, src, dst, directed, multi, count, flats):
    return {"allow_multiple_edge": multi, "directed": directed, "dst": dst, "properties": {},
            "src": src, "statistics": {"count": count},
            "store_info": {"flats": flats, "partition_counts": [], "partitioned_flats": []},
            "type": type_}


def write_schema(d, name, gtype, vertices, edges):
    with open(os.path.join(d, "schema.json"), "w") as f:
        json.dump({"edge": edges, "graph_name": name, "graph_type": gtype, "vertex": vertices},
                  f, indent=4, sort_keys=True)
        f.write("\n")


def main():
    d = os.path.join(ROOT, "knows")
    os.makedirs(d, exist_ok=True)
    diskarray(os.path.join(d, "Person-knows-Person_csr_src_dst_index"), "q", [0, 3, 6, 8, 10, 10])
    diskarray(os.path.join(d, "Person-knows-Person_csr_src_dst"), "I", [1, 1, 2, 0, 0, 2, 0, 1, 3, 3])
    diskarray(os.path.join(d, "Person-knows-Person_src_dst"), "II",
              [(0, 1), (0, 1), (0, 2), (1, 0), (1, 0), (1, 2), (2, 0), (2, 1), (3, 3), (3, 3)])
    diskarray(os.path.join(d, "Person.id"), "I", [0, 1, 2, 3, 4])
    write_schema(d, "", "Social Network", [vertex("Person", 5)],
                 [edge("knows", "Person", "Person", False, False, 10, [1])])

    d = os.path.join(ROOT, "fin")
    os.makedirs(d, exist_ok=True)
    diskarray(os.path.join(d, "Account-transfer->Account_csr_src_dst_index"), "q", [0, 3, 4, 5])
    diskarray(os.path.join(d, "Account-transfer->Account_csr_src_dst"), "I", [1, 1, 1, 2, 0])
    diskarray(os.path.join(d, "Account<-transfer-Account_csr_dst_src_index"), "q", [0, 1, 4, 5])
    diskarray(os.path.join(d, "Account<-transfer-Account_csr_dst_src"), "I", [2, 0, 0, 0, 1])
    diskarray(os.path.join(d, "Medium-signIn->Account_csr_src_dst_index"), "q", [0, 1, 3])
    diskarray(os.path.join(d, "Medium-signIn->Account_csr_src_dst"), "I", [2, 0, 2])
    diskarray(os.path.join(d, "Account<-signIn-Medium_csr_dst_src_index"), "q", [0, 1, 1, 3])
    diskarray(os.path.join(d, "Account<-signIn-Medium_csr_dst_src"), "I", [1, 0, 1])
    diskarray(os.path.join(d, "Account.id"), "I", [0, 1, 2])
    diskarray(os.path.join(d, "Medium.id"), "I", [0, 1])
    write_schema(d, "", "Bank Transfer", [vertex("Account", 3), vertex("Medium", 2)],
                 [edge("transfer", "Account", "Account", True, True, 5, []),
                  edge("signIn", "Medium", "Account", True, False, 3, [])])


if __name__ == "__main__":
    main()
