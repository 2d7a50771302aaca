//! The crate's API on a device: objval is the cut of the returned partition,
//! a seed fixes the result, the option array round-trips, orderings are
//! permutations, and .graph files round-trip.  Skipped without a device.
use rmetis::*;

fn grid(rows: usize, cols: usize) -> Graph {
    let mut xadj = vec![0 as Idx];
    let mut adjncy = Vec::new();
    for r in 0..rows {
        for c in 0..cols {
            if r > 0 { adjncy.push(((r - 1) * cols + c) as Idx); }
            if c > 0 { adjncy.push((r * cols + c - 1) as Idx); }
            if c + 1 < cols { adjncy.push((r * cols + c + 1) as Idx); }
            if r + 1 < rows { adjncy.push(((r + 1) * cols + c) as Idx); }
            xadj.push(adjncy.len() as Idx);
        }
    }
    Graph::new_unweighted(rows * cols, xadj, adjncy).unwrap()
}

#[test]
fn options_round_trip() {
    let o = Options { rtype: RType::Greedy, seed: 7, ncuts: 3, ..Options::for_kway() };
    let mut raw = o.to_raw();
    assert_eq!(raw[METIS_OPTION_PTYPE], -1);
    raw[METIS_OPTION_PTYPE] = 0;
    let p = Options::from_raw(&raw);
    assert_eq!((p.rtype, p.seed, p.ncuts, p.ufactor), (RType::Greedy, 7, 3, 30));
}

#[test]
fn invalid_graphs_are_input_errors() {
    // a self-loop (graph/validate.rs)
    assert!(matches!(Graph::new_unweighted(2, vec![0, 1, 2], vec![0, 0]), Err(MetisError::Input(_))));
}

#[test]
fn kway_recursive_and_nd() {
    if device_count() < 1 {
        return;
    }
    let g = grid(32, 32);
    for rtype in [RType::Greedy, RType::Fm] {
        let o = Options { rtype, seed: 1, ..Options::for_kway() };
        let r = part_graph_kway(&g, 8, None, None, &o).unwrap();
        assert_eq!(r.objval, g.edge_cut(&r.part).unwrap());
        let again = part_graph_kway(&g, 8, None, None, &o).unwrap();
        assert_eq!(r.part, again.part);
        let rb = part_graph_recursive(&g, 8, None, None, &o).unwrap();
        assert_eq!(rb.objval, g.edge_cut(&rb.part).unwrap());
        assert!(rb.part.iter().all(|&p| (0..8).contains(&p)));
    }
    let nd = node_nd(&g, &Options { seed: 1, ..Options::default() }).unwrap();
    let mut seen = vec![false; g.nvtxs];
    for v in 0..g.nvtxs {
        seen[nd.perm[v] as usize] = true;
        assert_eq!(nd.iperm[nd.perm[v] as usize] as usize, v);
    }
    assert!(seen.into_iter().all(|s| s));
    assert!(matches!(part_graph_kway(&g, 1, None, None, &Options::default()), Err(MetisError::Input(_))));
}

#[test]
fn graph_files_round_trip() {
    let g = grid(5, 7);
    let path = std::env::temp_dir().join("rmetis_hip_api_test.graph");
    let p = path.to_str().unwrap();
    write_graph(p, &g).unwrap();
    let h = read_graph(p).unwrap();
    assert_eq!((h.nvtxs, h.xadj.clone(), h.adjncy.clone()), (g.nvtxs, g.xadj.clone(), g.adjncy.clone()));
    let _ = std::fs::remove_file(path);
}
