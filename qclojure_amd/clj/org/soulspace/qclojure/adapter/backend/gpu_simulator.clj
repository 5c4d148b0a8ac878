(ns org.soulspace.qclojure.adapter.backend.gpu-simulator
  "QuantumBackend and BatchJobBackend over libqsv, the MI355X state-vector
  engine, called through the JDK Foreign Function & Memory API (JDK >= 22).

  Drop-in replacement for LocalQuantumSimulator
  (src/org/soulspace/qclojure/adapter/backend/ideal_simulator.clj:100-176):
  same protocols (application/backend.clj:72-165), same job lifecycle and
  result map (ideal_simulator.clj:70-96, circuit.clj:1783-1791), but the
  amplitudes live in HBM and every gate and every result reduction runs in a
  HIP kernel. Circuits are decoded on the host into 96-byte qsv_op records
  exactly as apply-gate-to-state dispatches them (circuit.clj:889-1072).

  Library path: the system property qsv.library (default \"libqsv.so\" on the
  loader path). Config keys of create-simulator:
    :max-qubits        largest circuit accepted (default 34)
    :precision         :c128 (default) or :c64
    :seed              base seed of shots and mid-circuit measurements
    :workers           jobs run concurrently (default 4)
    :msb-swap          true: SWAP/iSWAP on MSB positions (the reference's
                       shifts use LSB positions, gate.clj:768-778, 823-833)
    :fuse              true (default): fused LDS-tiled passes
    :keep-final-state  true (default): :final-state stays in HBM until read
                       or until cleanup-completed-jobs! drops the job
    :shadow-max-qubits complex64 jobs up to this size (default 24) also run
                       a complex128 shadow and report fidelity against it"
  (:require [fastmath.complex :as fc]
            [org.soulspace.qclojure.application.backend :as backend]
            [org.soulspace.qclojure.domain.circuit :as circuit]
            [org.soulspace.qclojure.domain.hamiltonian :as ham]
            [org.soulspace.qclojure.domain.operation-registry :as opreg])
  (:import [java.lang.foreign Arena FunctionDescriptor Linker Linker$Option MemoryLayout
            MemorySegment SymbolLookup ValueLayout ValueLayout$OfDouble ValueLayout$OfInt
            ValueLayout$OfLong AddressLayout]
           [java.lang.invoke MethodHandle]
           [java.util.concurrent ExecutorService Executors]))

;;;
;;; The C ABI (include/qsv.h)
;;;
(def ^:private ^ValueLayout$OfInt I ValueLayout/JAVA_INT)
(def ^:private ^ValueLayout$OfLong J ValueLayout/JAVA_LONG)
(def ^:private ^ValueLayout$OfDouble D ValueLayout/JAVA_DOUBLE)
(def ^:private ^AddressLayout P ValueLayout/ADDRESS)

(def ^:private ^Linker linker (Linker/nativeLinker))

(def ^:private ^SymbolLookup lib
  (delay (SymbolLookup/libraryLookup ^String (System/getProperty "qsv.library" "libqsv.so")
                                     (Arena/global))))

(defn- down
  "Downcall handle of `sym`; ret nil for void."
  ^MethodHandle [^String sym ret & args]
  (let [addr (.orElseThrow (.find ^SymbolLookup @lib sym))
        fd (if ret
             (FunctionDescriptor/of ^MemoryLayout ret ^"[Ljava.lang.foreign.MemoryLayout;" (into-array MemoryLayout args))
             (FunctionDescriptor/ofVoid ^"[Ljava.lang.foreign.MemoryLayout;" (into-array MemoryLayout args)))]
    (.downcallHandle linker ^MemorySegment addr ^FunctionDescriptor fd
                     ^"[Ljava.lang.foreign.Linker$Option;" (into-array Linker$Option []))))

(defmacro ^:private defdown [name sym ret & args]
  `(def ^:private ~name (delay (down ~sym ~ret ~@args))))

(defdown qsv-init           "qsv_init"           I I)
(defdown qsv-last-error     "qsv_last_error"     P)
(defdown qsv-version        "qsv_version"        P)
(defdown qsv-device-query   "qsv_device_query"   I P)
(defdown qsv-state-create   "qsv_state_create"   I I I P)
(defdown qsv-state-destroy  "qsv_state_destroy"  I P)
(defdown qsv-state-upload   "qsv_state_upload"   I P P J J)
(defdown qsv-state-download "qsv_state_download" I P P J J)
(defdown qsv-state-sync     "qsv_state_sync"     I P)
(defdown qsv-apply-program  "qsv_apply_program"  I P P J I J)
(defdown qsv-norm2          "qsv_norm2"          I P P)
(defdown qsv-probabilities  "qsv_probabilities"  I P J J P)
(defdown qsv-sample         "qsv_sample"         I P J J P)
(defdown qsv-expect-pauli   "qsv_expect_pauli"   I P P J P)
(defdown qsv-expect-1q      "qsv_expect_1q"      I P I P P)
(defdown qsv-inner          "qsv_inner"          I P P P)
(defdown qsv-state-compare  "qsv_state_compare"  I P P P)
(defdown qsv-batch-expect-z "qsv_batch_expect_z" I P P)
(defdown qsv-mat-create     "qsv_mat_create"     I J J P)
(defdown qsv-mat-destroy    "qsv_mat_destroy"    I P)
(defdown qsv-mat-upload     "qsv_mat_upload"     I P P)
(defdown qsv-mat-view-state "qsv_mat_view_state" I P P)
(defdown qsv-la-gemv        "qsv_la_gemv"        I P P P)
(defdown qsv-la-dot         "qsv_la_dot"         I I P P P)

(def ^:private qsv-c128 0)
(def ^:private qsv-c64 1)
(def ^:private flag-msb-swap 0x1)
(def ^:private flag-fuse 0x2)

(defn- call ^Object [h & args]
  (.invokeWithArguments ^MethodHandle @h ^java.util.List (vec args)))

(defn- c-string [^MemorySegment p]
  (if (= (.address p) 0) "" (.getString (.reinterpret p 4096) 0)))

(defn- check
  "Turn a non-zero libqsv status into ex-info carrying qsv_last_error."
  [rc]
  (let [rc (int rc)]
    (when-not (zero? rc)
      (throw (ex-info (c-string (call qsv-last-error)) {:qsv-code rc})))))

;;;
;;; Decoding circuit operations into qsv_op records
;;; (apply-gate-to-state, circuit.clj:933-1072; apply-general-controlled-gate,
;;; circuit.clj:889-930). Kind numbers are enum qsv_op_kind.
;;;
(def ^:private op-bytes 96)

(def ^:private one-qubit-kind
  {:x 1 :y 2 :z 3 :h 4 :s 5 :s-dag 6 :t 7 :t-dag 8 :rx 9 :ry 10 :rz 11 :phase 12})

(def ^:private needs-angle #{:rx :ry :rz :phase})

;; [kind needs-angle? message] of the two-qubit control/target gates the
;; dispatcher accepts. :cy is not among them: the dispatcher has no :cy case
;; and fails it as an unknown gate type; CY (kind 15) is reached only through
;; :controlled with :target-gate :y.
(def ^:private control-target
  {:cnot           [13 false "CNOT requires both control and target qubits"]
   :cz             [14 false "CZ requires both control and target qubits"]
   :crx            [16 true  "CRX requires control, target qubits and angle"]
   :cry            [17 true  "CRY requires control, target qubits and angle"]
   :crz            [18 true  "CRZ requires control, target qubits and angle"]
   :rydberg-cz     [23 false "Rydberg CZ requires both control and target qubits"]
   :rydberg-cphase [24 true  "Rydberg CPhase requires control, target qubits and phase angle"]})

(def ^:private global-kind
  {:global-rx 26 :global-ry 27 :global-rz 28 :global-h 29 :global-x 30 :global-y 31 :global-z 32})

(def ^:private kind-measure 33)

(defn- qubit! [n what q]
  (when-not (and (integer? q) (<= 0 q) (< q n))
    (throw (ex-info (str what " must be a qubit index below " n) {what q})))
  (long q))

(defn- qubit-mask! [n what qs]
  (when (or (empty? qs) (not= (count qs) (count (distinct qs))))
    (throw (ex-info (str what ": qubits must be a non-empty list of distinct indices") {what qs})))
  (reduce (fn [m q] (bit-or m (bit-shift-left 1 (qubit! n what q)))) 0 qs))

(defn- rec
  "One decoded record: [kind q0 q1 q2 mask angle]."
  ([kind q0] (rec kind q0 0 0 0 0.0))
  ([kind q0 q1] (rec kind q0 q1 0 0 0.0))
  ([kind q0 q1 q2 mask angle] [kind q0 q1 q2 mask (double angle)]))

(defn- distinct! [msg & qs]
  (when-not (apply distinct? qs) (throw (ex-info msg {:qubits (vec qs)}))))

(defn- decode-controlled
  "apply-general-controlled-gate (circuit.clj:889-930): one control and one
  target lower to CNOT/CZ/CY; :h :s :t act on the target unconditionally;
  other target gates leave the state unchanged; more controls fail."
  [n {:keys [control-qubits target-gate target-qubits] :as params}]
  (when-not (and (seq control-qubits) target-gate)
    (throw (ex-info "Controlled gate requires control-qubits and target-gate" {:params params})))
  (let [targets (or target-qubits [0])]
    (if (and (= 1 (count control-qubits)) (= 1 (count targets)))
      (let [c (qubit! n :control (first control-qubits))
            t (qubit! n :target (first targets))]
        (case target-gate
          :x (do (distinct! "control and target must differ" c t) [(rec 13 c t)])
          :z (do (distinct! "control and target must differ" c t) [(rec 14 c t)])
          :y (do (distinct! "control and target must differ" c t) [(rec 15 c t)])
          :h [(rec 4 t)]
          :s [(rec 5 t)]
          :t [(rec 7 t)]
          []))
      (throw (ex-info "Multi-controlled gates with >1 control qubit not fully implemented"
                      {:control-qubits control-qubits :target-gate target-gate :target-qubits targets})))))

(defn decode-operation
  "Records of one operation map, or ex-info with the dispatcher's message."
  [n {:keys [operation-type operation-params] :as op}]
  (let [{:keys [target control control1 control2 target1 target2 qubit1 qubit2 angle
                qubit-indices measurement-qubits]} operation-params]
    (if (= operation-type :measure)
      (if (seq measurement-qubits)
        [(rec kind-measure 0 0 0 (qubit-mask! n :measurement-qubits measurement-qubits) 0.0)]
        (throw (ex-info "Measure requires measurement-qubits parameter" {:operation op})))
      (let [g (opreg/resolve-gate-alias operation-type)]
        (cond
          (one-qubit-kind g)
          (do (when (and (needs-angle g) (nil? angle))
                (throw (ex-info (str (name g) " requires an angle") {:gate op})))
              [(rec (one-qubit-kind g) (qubit! n :target (or target 0)) 0 0 0 (or angle 0.0))])

          (control-target g)
          (let [[k angled? msg] (control-target g)]
            (when (or (nil? control) (nil? target) (and angled? (nil? angle)))
              (throw (ex-info msg {:gate op})))
            (let [c (qubit! n :control control) t (qubit! n :target target)]
              (distinct! "control and target must differ" c t)
              [(rec k c t 0 0 (or angle 0.0))]))

          (#{:swap :iswap} g)
          (do (when (or (nil? qubit1) (nil? qubit2))
                (throw (ex-info (str (if (= g :swap) "SWAP" "iSWAP")
                                     " requires both qubit1 and qubit2 parameters") {:gate op})))
              (let [a (qubit! n :qubit1 qubit1) b (qubit! n :qubit2 qubit2)]
                (distinct! "qubit1 and qubit2 must differ" a b)
                [(rec (if (= g :swap) 19 20) a b)]))

          (= g :toffoli)
          (do (when (some nil? [control1 control2 target])
                (throw (ex-info "Toffoli requires control1, control2, and target parameters" {:gate op})))
              (let [qs (mapv #(qubit! n %1 %2) [:control1 :control2 :target] [control1 control2 target])]
                (apply distinct! "Toffoli qubits must be distinct" qs)
                [(rec 21 (qs 0) (qs 1) (qs 2) 0 0.0)]))

          (= g :fredkin)
          (do (when (some nil? [control target1 target2])
                (throw (ex-info "Fredkin requires control, target1, and target2 parameters" {:gate op})))
              (let [qs (mapv #(qubit! n %1 %2) [:control :target1 :target2] [control target1 target2])]
                (apply distinct! "Fredkin qubits must be distinct" qs)
                [(rec 22 (qs 0) (qs 1) (qs 2) 0 0.0)]))

          (= g :rydberg-blockade)
          (do (when (or (empty? qubit-indices) (nil? angle))
                (throw (ex-info "Rydberg blockade requires qubit-indices vector and phase angle" {:gate op})))
              [(rec 25 0 0 0 (qubit-mask! n :qubit-indices qubit-indices) angle)])

          (global-kind g)
          (do (when (and (#{:global-rx :global-ry :global-rz} g) (nil? angle))
                (throw (ex-info (str "Global " (subs (.toUpperCase (name g)) 7) " requires rotation angle")
                                {:gate op})))
              [(rec (global-kind g) 0 0 0 0 (or angle 0.0))])

          (= g :controlled)
          (decode-controlled n operation-params)

          :else
          (throw (ex-info "Unknown gate type" {:operation-type g})))))))

(defn- encode-program
  "All records of `circuit` packed into qsv_op structs in `arena`: [segment count]."
  [^Arena arena circuit]
  (let [n (:num-qubits circuit)
        recs (into [] (mapcat #(decode-operation n %)) (:operations circuit))
        ^MemorySegment seg (.allocate arena (long (max op-bytes (* op-bytes (count recs)))) 16)]
    (doseq [[i [kind q0 q1 q2 mask angle]] (map-indexed vector recs)]
      (let [base (long (* i op-bytes))]
        (.set seg I base (int kind))
        (.set seg I (+ base 4) (int q0))
        (.set seg I (+ base 8) (int q1))
        (.set seg I (+ base 12) (int q2))
        (.set seg J (+ base 16) (long mask))
        (.set seg D (+ base 24) (double angle))))
    [seg (count recs)]))

;;;
;;; Device states
;;;
(defn- state-create ^MemorySegment [n precision]
  (with-open [a (Arena/ofConfined)]
    (let [out (.allocate a P)]
      (check (call qsv-state-create (int n) (int precision) out))
      (.get ^MemorySegment out P 0))))

(defn- state-destroy! [^MemorySegment st]
  (when st (call qsv-state-destroy st)))

(defn- upload-state!
  "Amplitudes of a reference state map ({:state-vector [Vec2 ...]}) into st."
  [st {:keys [state-vector]}]
  (with-open [a (Arena/ofConfined)]
    (let [cnt (count state-vector)
          ^MemorySegment buf (.allocate a (long (* 16 cnt)) 16)]
      (doseq [[i z] (map-indexed vector state-vector)]
        (.setAtIndex buf D (long (* 2 i)) (double (fc/re z)))
        (.setAtIndex buf D (long (inc (* 2 i))) (double (fc/im z))))
      (check (call qsv-state-upload st buf 0 (long cnt))))))

(defn- download-amplitudes
  "Amplitudes [offset, offset+count) as a vector of fastmath complex numbers."
  [st offset cnt]
  (with-open [a (Arena/ofConfined)]
    (let [^MemorySegment buf (.allocate a (long (* 16 (max 1 cnt))) 16)]
      (check (call qsv-state-download st buf (long offset) (long cnt)))
      (mapv (fn [i] (fc/complex (.getAtIndex buf D (long (* 2 i))) (.getAtIndex buf D (long (inc (* 2 i))))))
            (range cnt)))))

(defn- probabilities [st offset cnt]
  (with-open [a (Arena/ofConfined)]
    (let [^MemorySegment buf (.allocate a (long (* 8 (max 1 cnt))) 8)]
      (check (call qsv-probabilities st (long offset) (long cnt) buf))
      (vec (.toArray buf D)))))

(defn- norm2 ^double [st]
  (with-open [a (Arena/ofConfined)]
    (let [out (.allocate a D)]
      (check (call qsv-norm2 st out))
      (.get ^MemorySegment out D 0))))

(defn- matrix->doubles
  "A 2x2 matrix of fastmath complex numbers as the row-major (re, im) doubles of qsv_expect_1q."
  [^Arena a m]
  (let [^MemorySegment seg (.allocate a 64 8)]
    (doseq [[k z] (map-indexed vector (for [row m z row] z))]
      (.setAtIndex seg D (long (* 2 k)) (double (fc/re z)))
      (.setAtIndex seg D (long (inc (* 2 k))) (double (fc/im z))))
    seg))

(defn- expect-1q
  "<psi| I..O_q..I |psi> on the device, as a fastmath complex number."
  [st q observable]
  (with-open [a (Arena/ofConfined)]
    (let [out (.allocate a 16 8)]
      (check (call qsv-expect-1q st (int q) (matrix->doubles a observable) out))
      (fc/complex (.getAtIndex ^MemorySegment out D 0) (.getAtIndex ^MemorySegment out D 1)))))

(defn- expect-dense
  "<psi| O psi> for an observable of the state's full dimension: the device
  GEMV and inner product of the :hip math backend on a view of the state
  (complex128 states only)."
  [st observable]
  (with-open [a (Arena/ofConfined)]
    (let [dim (count observable)
          pm (.allocate a P) po (.allocate a P) py (.allocate a P)
          host (.allocate a (long (* 16 dim dim)) 16)
          out (.allocate a 16 8)]
      (doseq [[k z] (map-indexed vector (for [row observable z row] z))]
        (.setAtIndex ^MemorySegment host D (long (* 2 k)) (double (fc/re z)))
        (.setAtIndex ^MemorySegment host D (long (inc (* 2 k))) (double (fc/im z))))
      (check (call qsv-mat-view-state st pm))
      (let [psi (.get ^MemorySegment pm P 0)]
        (try
          (check (call qsv-mat-create (long dim) (long dim) po))
          (let [o (.get ^MemorySegment po P 0)]
            (try
              (check (call qsv-mat-upload o host))
              (check (call qsv-mat-create (long dim) 1 py))
              (let [y (.get ^MemorySegment py P 0)]
                (try
                  (check (call qsv-la-gemv o psi y))
                  (check (call qsv-la-dot (int 1) psi y out))
                  (fc/complex (.getAtIndex ^MemorySegment out D 0) (.getAtIndex ^MemorySegment out D 1))
                  (finally (call qsv-mat-destroy y))))
              (finally (call qsv-mat-destroy o))))
          (finally (call qsv-mat-destroy psi)))))))

(defn- expectation [st n observable target]
  (cond
    (some? target) (expect-1q st target observable)
    (= n 1) (expect-1q st 0 observable)
    (= (count observable) (bit-shift-left 1 n)) (expect-dense st observable)
    :else (throw (ex-info "observable dimension does not match the state (give a target qubit for a 2x2 observable)"
                          {:dimension (count observable) :num-qubits n}))))

(defn- mat-square [m]
  (let [k (count m)]
    (mapv (fn [i] (mapv (fn [j] (reduce fc/add (fc/complex 0.0 0.0)
                                        (map #(fc/mult (get-in m [i %]) (get-in m [% j])) (range k))))
                        (range k)))
          (range k))))

;;;
;;; Result extraction on the device (result.clj:179-457, 535-639)
;;;
(defn- measurement-results
  "extract-measurement-results, ideal branch (result.clj:179-252): shots full-register draws."
  [st n {:keys [shots qubits]} seed]
  (let [shots (long (or shots 1))
        outcomes (with-open [a (Arena/ofConfined)]
                   (let [^MemorySegment buf (.allocate a (long (* 8 (max 1 shots))) 8)]
                     (check (call qsv-sample st shots (long seed) buf))
                     (mapv #(.getAtIndex buf J (long %)) (range shots))))
        freq (frequencies outcomes)]
    {:measurement-outcomes outcomes
     :measurement-probabilities (delay (probabilities st 0 (bit-shift-left 1 n)))
     :empirical-probabilities (into {} (map (fn [[o c]] [o (/ c shots)]) freq))
     :shot-count shots
     :measurement-qubits (or qubits (vec (range n)))
     :frequencies freq
     :source :ideal-simulation}))

(defn- hamiltonian-result
  "extract-hamiltonian-expectation (result.clj:327-344): one device reduction
  per batch of Pauli strings, then the term-order FP64 sum of
  hamiltonian-expectation (hamiltonian.clj:96-114)."
  [st h]
  (when-not (ham/validate-hamiltonian h)
    (throw (ex-info "invalid hamiltonian" {:hamiltonian h})))
  (let [terms (vec h)
        values (if (empty? terms)
                 []
                 (with-open [a (Arena/ofConfined)]
                   (let [strings (.allocateFrom a ^String (apply str (map :pauli-string terms)))
                         ^MemorySegment out (.allocate a (long (* 8 (count terms))) 8)]
                     (check (call qsv-expect-pauli st strings (long (count terms)) out))
                     (mapv #(.getAtIndex out D (long %)) (range (count terms))))))]
    {:energy-expectation (reduce + (map (fn [t v] (* (:coefficient t) v)) terms values))
     :hamiltonian h
     :measurement-groups (ham/group-commuting-terms h)
     :measurement-bases (ham/group-pauli-terms-by-measurement-basis h)}))

(defn- fidelity-results
  "extract-fidelity-result (result.clj:437-457): |<psi|phi>| per reference
  state (the modulus, as state-fidelity returns it)."
  [st n precision references]
  {:fidelities
   (into {}
         (map-indexed
          (fn [k ref]
            (let [tmp (state-create n precision)]
              (try
                (upload-state! tmp ref)
                (with-open [a (Arena/ofConfined)]
                  (let [out (.allocate a 16 8)]
                    (check (call qsv-inner st tmp out))
                    [(str "reference-" k)
                     (Math/hypot (.getAtIndex ^MemorySegment out D 0) (.getAtIndex ^MemorySegment out D 1))]))
                (finally (state-destroy! tmp)))))
          references))
   :reference-states references})

(defn- bits->index [bits]
  (reduce (fn [acc b] (+ (* 2 acc) (long b))) 0 bits))

(defn- basis-label [i n]
  (let [s (Long/toBinaryString i)]
    (str "|" (apply str (repeat (- n (count s)) \0)) s "⟩")))

(defn- extract-results
  "extract-results (result.clj:535-639) with every amplitude read on the device."
  [base specs st n precision seed]
  (cond-> base
    (:measurements specs)
    (assoc :measurement-results (measurement-results st n (:measurements specs) seed))

    (:expectation specs)
    (assoc :expectation-results
           (let [{:keys [observables targets]} (:expectation specs)]
             (mapv (fn [o t] {:expectation-value (fc/re (expectation st n o t)) :observable o
                              :target-qubits (when (some? t) [t])})
                   observables (or targets (repeat nil)))))

    (:variance specs)
    (assoc :variance-results
           (let [{:keys [observables targets]} (:variance specs)]
             (mapv (fn [o t]
                     (let [e (fc/re (expectation st n o t))
                           e2 (fc/re (expectation st n (mat-square o) t))
                           v (- e2 (* e e))]
                       {:variance-value v :standard-deviation (if (neg? v) Double/NaN (Math/sqrt v))
                        :observable o :target-qubits (when (some? t) [t])}))
                   observables (or targets (repeat nil)))))

    (:hamiltonian specs)
    (assoc :hamiltonian-result (hamiltonian-result st (:hamiltonian specs)))

    (:probabilities specs)
    (assoc :probability-results
           (let [{:keys [qubits targets]} (:probabilities specs)]
             (if targets
               {:probability-outcomes (into {} (map (fn [t] [t (first (probabilities st (if (coll? t) (bits->index t) t) 1))])
                                                    targets))
                :target-states targets :target-qubits qubits}
               (let [allp (probabilities st 0 (bit-shift-left 1 n))]
                 {:probability-outcomes (into {} (map-indexed vector allp))
                  :target-qubits (or qubits (vec (range n))) :all-probabilities allp}))))

    (:amplitudes specs)
    (assoc :amplitude-results
           (let [bs (:basis-states (:amplitudes specs))]
             {:amplitude-values (into {} (map (fn [b] [b (first (download-amplitudes st b 1))]) bs))
              :basis-states bs}))

    (:state-vector specs)
    (assoc :state-vector-result {:state-vector (delay (download-amplitudes st 0 (bit-shift-left 1 n)))
                                 :num-qubits n
                                 :basis-labels (delay (mapv #(basis-label % n) (range (bit-shift-left 1 n))))})

    (:fidelity specs)
    (assoc :fidelity-results (fidelity-results st n precision (:references (:fidelity specs))))))

;;;
;;; Precision report of complex64 results: |<psi32|psi64>|^2 and max |d amplitude|
;;; against a complex128 shadow run, and the norm drift |<psi32|psi32> - 1|
;;;
(defn- precision-report [st circuit init flags seed precision shadow-max]
  (let [n (:num-qubits circuit)]
    (if (not= precision qsv-c64)
      {:precision :c128}
      (let [drift (Math/abs (- (norm2 st) 1.0))]
        (if (> n shadow-max)
          {:precision :c64 :norm-drift drift}
          (let [shadow (state-create n qsv-c128)]
            (try
              (when init (upload-state! shadow init))
              (with-open [a (Arena/ofConfined)]
                (let [[seg cnt] (encode-program a circuit)
                      ^MemorySegment out (.allocate a 40 8)]
                  (check (call qsv-apply-program shadow seg (long cnt) (int flags) (long seed)))
                  (check (call qsv-state-compare st shadow out))
                  (let [re (.getAtIndex out D 0) im (.getAtIndex out D 1)]
                    {:precision :c64 :norm-drift drift
                     :fidelity-vs-c128 (+ (* re re) (* im im))
                     :max-abs-amplitude-error (.getAtIndex out D 4)
                     :reference :complex128-device-shadow})))
              (finally (state-destroy! shadow)))))))))

;;;
;;; Jobs (ideal_simulator.clj:56-96, 120-176)
;;;
(defonce ^:private simulator-state (atom {:job-counter 0 :active-jobs {} :batches {}}))

(defn- next-id [prefix]
  (let [k (:job-counter (swap! simulator-state update :job-counter inc))]
    (str prefix "_" k "_" (System/currentTimeMillis))))

(defn- final-state-handle
  "The :final-state of a result: a map whose :state-vector is realised from
  HBM on first deref; :release frees the device state."
  [st n]
  (let [handle (atom st)]
    {:num-qubits n
     :state-vector (delay (if-let [s @handle] (download-amplitudes s 0 (bit-shift-left 1 n))
                            (throw (ex-info "final state released" {}))))
     :release (fn [] (when-let [s @handle] (when (compare-and-set! handle s nil) (state-destroy! s))))}))

(defn- run-circuit
  "The private execute-circuit of ideal_simulator.clj:70-96 on the device."
  [config circuit options seed]
  (let [start (System/currentTimeMillis)
        n (:num-qubits circuit)
        precision (if (= :c64 (:precision config)) qsv-c64 qsv-c128)
        flags (bit-or (if (:msb-swap config) flag-msb-swap 0)
                      (if (false? (:fuse config)) 0 flag-fuse))
        max-qubits (get config :max-qubits 34)]
    (try
      (when (> n max-qubits)
        (throw (ex-info (str "circuit has " n " qubits, backend max-qubits is " max-qubits) {})))
      (let [init (:initial-state options)
            _ (when (and init (not= (:num-qubits init) n))
                (throw (ex-info "initial state and circuit have different numbers of qubits" {})))
            st (state-create n precision)]
        (try
          (when init (upload-state! st init))
          (with-open [a (Arena/ofConfined)]
            (let [[seg cnt] (encode-program a circuit)]
              (check (call qsv-apply-program st seg (long cnt) (int flags) (long seed)))))
          (let [specs (:result-specs options)
                fs (final-state-handle st n)
                base {:final-state fs
                      :result-types (set (keys specs))
                      :circuit circuit
                      :circuit-metadata {:circuit-depth (circuit/circuit-depth circuit)
                                         :circuit-operation-count (circuit/circuit-operation-count circuit)
                                         :circuit-gate-count (circuit/circuit-gate-count circuit)}
                      :precision-report (precision-report st circuit init flags seed precision
                                                          (get config :shadow-max-qubits 24))}
                results (if (seq specs)
                          (extract-results base specs st n precision (bit-xor seed (unchecked-long 0xA5A5A5A5A5A5A5A5)))
                          base)]
            (check (call qsv-state-sync st))
            (when (false? (:keep-final-state config))
              (doseq [v (vals results)] (when (map? v) (doseq [x (vals v)] (when (delay? x) (force x)))))
              (force (:state-vector fs))
              ((:release fs)))
            {:job-status :completed
             :results results
             :execution-time-ms (- (System/currentTimeMillis) start)})
          (catch Throwable t
            (state-destroy! st)
            (throw t))))
      (catch Exception e
        {:job-status :failed
         :error-message (.getMessage e)
         :exception-type (.getName (class e))}))))

(defn- release-result! [result]
  (when-let [release (get-in result [:results :final-state :release])] (release)))

(defn- job-finished!
  "Store a job's outcome unless the job was cancelled meanwhile (the
  reference overwrites a cancelled job with its result)."
  [job-id result]
  (let [[old _] (swap-vals! simulator-state update-in [:active-jobs job-id]
                            (fn [job] (if (and job (= :running (:status job)))
                                        (assoc job :status (:job-status result) :result result
                                               :completed-at (System/currentTimeMillis))
                                        job)))]
    (when-not (= :running (get-in old [:active-jobs job-id :status]))
      (release-result! result))))

(defrecord GpuSimulator [config ^ExecutorService pool seeds]
  backend/QuantumBackend
  (backend-info [this]
    {:backend-type :simulator
     :backend-name "MI355X GPU State-Vector Simulator"
     :description "State-vector simulator on hand-written CDNA4 HIP kernels (libqsv)"
     :backend-config config
     :max-qubits (get config :max-qubits 34)
     :capabilities #{:quantum-backend :batch-job-backend}
     :device (backend/device this)
     :version (c-string (call qsv-version))})

  (device [_]
    {:id :mi355x-simulator
     :name "MI355X State-Vector Simulator"
     :provider "QClojure"
     :platform "Local"
     :technology :simulator
     :num-qubits (get config :max-qubits 34)
     :topology :all-to-all
     :connectivity :full
     :native-gates opreg/native-simulator-gate-set
     :virtual-gates #{}
     :supported-operations opreg/native-simulator-gate-set
     :measurement-basis :any
     :noise-model {}
     :performance {:gate-fidelity 1.0 :readout-fidelity 1.0 :gate-times {:all 0}
                   :simulation-time {:exponential-scaling true}}})

  (available? [_]
    (try (zero? (int (call qsv-init (int (get config :device 0))))) (catch Throwable _ false)))

  (submit-circuit [_ circuit options]
    (let [job-id (next-id "gpu_job")
          seed (unchecked-add (long (get config :seed 0x5EED)) (long (swap! seeds inc)))]
      (swap! simulator-state assoc-in [:active-jobs job-id]
             {:job-id job-id :circuit circuit :options options :status :queued :result nil
              :created-at (System/currentTimeMillis) :completed-at nil})
      (.submit pool ^Runnable
               (fn []
                 (let [[old _] (swap-vals! simulator-state update-in [:active-jobs job-id]
                                           #(if (= :queued (:status %)) (assoc % :status :running) %))]
                   (when (= :queued (get-in old [:active-jobs job-id :status]))
                     (job-finished! job-id (run-circuit config circuit options seed))))))
      job-id))

  (job-status [_ job-id]
    (get-in @simulator-state [:active-jobs job-id :status] :not-found))

  (job-result [_ job-id]
    (if-let [job (get-in @simulator-state [:active-jobs job-id])]
      (if (#{:completed :failed} (:status job))
        (assoc (:result job) :job-id job-id)
        {:job-id job-id :job-status (:status job) :error-message "Job not completed"})
      {:job-id job-id :job-status :not-found :error-message "Job not found"}))

  (cancel-job [_ job-id]
    (let [[old _] (swap-vals! simulator-state update-in [:active-jobs job-id]
                              #(if (#{:queued :running} (:status %))
                                 (assoc % :status :cancelled :completed-at (System/currentTimeMillis))
                                 %))]
      (if-let [job (get-in old [:active-jobs job-id])]
        (if (#{:queued :running} (:status job)) :cancelled :cannot-cancel)
        :not-found)))

  (queue-status [_]
    (let [jobs (vals (:active-jobs @simulator-state))
          cnt (fn [s] (count (filter #(= s (:status %)) jobs)))]
      {:total-jobs (count jobs) :queued (cnt :queued) :running (cnt :running) :completed (cnt :completed)
       :backend-load 0.0 :estimated-wait-time 0}))

  backend/BatchJobBackend
  (batch-submit [this circuits options]
    (let [batch-id (next-id "gpu_batch")
          ids (mapv #(backend/submit-circuit this % options) circuits)]
      (swap! simulator-state assoc-in [:batches batch-id] ids)
      batch-id))
  (batch-submit [this circuits _device options]
    (backend/batch-submit this circuits options))

  (batch-status [this batch-id]
    (if-let [ids (get-in @simulator-state [:batches batch-id])]
      (let [st (into {} (map (fn [id] [id (backend/job-status this id)]) ids))]
        {:batch-id batch-id
         :status (if (every? #{:completed :failed :cancelled} (vals st)) :completed :running)
         :job-statuses st})
      {:batch-id batch-id :status :not-found}))

  (batch-results [this batch-id]
    (into {} (map (fn [id] [id (backend/job-result this id)])
                  (get-in @simulator-state [:batches batch-id])))))

(defn create-simulator
  "A GpuSimulator on device (:device config, default 0)."
  ([] (create-simulator {}))
  ([config]
   {:pre [(map? config)]}
   (check (call qsv-init (int (get config :device 0))))
   (->GpuSimulator config (Executors/newFixedThreadPool (int (get config :workers 4))) (atom 0))))

(defn reset-simulator-state!
  "Drop every job (freeing their device states), as ideal_simulator.clj:200-206."
  []
  (let [[old _] (reset-vals! simulator-state {:job-counter 0 :active-jobs {} :batches {}})]
    (doseq [job (vals (:active-jobs old))] (release-result! (:result job)))))

(defn cleanup-completed-jobs!
  "cleanup-completed-jobs! (ideal_simulator.clj:243-263): remove completed
  jobs older than max-age-ms, freeing their final states' device memory.
  Returns the number removed."
  ([] (cleanup-completed-jobs! (* 60 60 1000)))
  ([max-age-ms]
   (let [cutoff (- (System/currentTimeMillis) max-age-ms)
         stale? (fn [job] (and (= :completed (:status job)) (<= (or (:completed-at job) 0) cutoff)))
         [old _] (swap-vals! simulator-state update :active-jobs
                             #(into {} (remove (comp stale? val)) %))
         dropped (filter stale? (vals (:active-jobs old)))]
     (doseq [job dropped] (release-result! (:result job)))
     (count dropped))))

;;;
;;; Many small circuits with Z-only Hamiltonians in one library call (the
;;; variational fan-out of optimization.clj:255,283 without a job per circuit)
;;;
(defn batch-energies
  "<H> of each circuit's final state from |0..0>, where each Hamiltonian has
  Z/I Pauli strings only: one qsv_batch_expect_z call. circuits-and-hams is
  a sequence of [circuit hamiltonian]; returns a vector of energies (FP64,
  summed in term order)."
  [circuits-and-hams precision]
  (with-open [a (Arena/ofConfined)]
    (let [items (vec circuits-and-hams)
          cnt (count items)
          decoded (mapv (fn [[c _]] (into [] (mapcat #(decode-operation (:num-qubits c) %)) (:operations c))) items)
          total-ops (reduce + (map count decoded))
          terms (mapv (fn [[_ h]] (vec h)) items)
          total-terms (reduce + (map count terms))
          ^MemorySegment nq (.allocate a (long (* 4 (max 1 cnt))) 8)
          ^MemorySegment ooff (.allocate a (long (* 8 (inc cnt))) 8)
          ^MemorySegment ops (.allocate a (long (* op-bytes (max 1 total-ops))) 16)
          ^MemorySegment toff (.allocate a (long (* 8 (inc cnt))) 8)
          ^MemorySegment zmask (.allocate a (long (* 8 (max 1 total-terms))) 8)
          ^MemorySegment coeff (.allocate a (long (* 8 (max 1 total-terms))) 8)
          ^MemorySegment desc (.allocate a 64 8)
          ^MemorySegment out (.allocate a (long (* 8 (max 1 cnt))) 8)]
      (loop [i 0 o 0 t 0]
        (.setAtIndex ooff J (long i) (long o))
        (.setAtIndex toff J (long i) (long t))
        (when (< i cnt)
          (let [[c _] (items i)]
            (.setAtIndex nq I (long i) (int (:num-qubits c)))
            (doseq [[j [kind q0 q1 q2 mask angle]] (map-indexed vector (decoded i))]
              (let [base (long (* (+ o j) op-bytes))]
                (.set ops I base (int kind))
                (.set ops I (+ base 4) (int q0))
                (.set ops I (+ base 8) (int q1))
                (.set ops I (+ base 12) (int q2))
                (.set ops J (+ base 16) (long mask))
                (.set ops D (+ base 24) (double angle))))
            (doseq [[j {:keys [coefficient pauli-string]}] (map-indexed vector (terms i))]
              (when-not (every? #{\I \Z} pauli-string)
                (throw (ex-info "batch-energies takes Z/I Pauli strings only" {:pauli-string pauli-string})))
              (.setAtIndex zmask J (long (+ t j))
                           (reduce (fn [m [q ch]] (if (= ch \Z) (bit-or m (bit-shift-left 1 q)) m))
                                   0 (map-indexed vector pauli-string)))
              (.setAtIndex coeff D (long (+ t j)) (double coefficient)))
            (recur (inc i) (+ o (count (decoded i))) (+ t (count (terms i)))))))
      ;; qsv_batch_desc: u64 n; 6 pointers; i32 precision
      (.set desc J 0 (long cnt))
      (doseq [[k seg] (map-indexed vector [nq ooff ops toff zmask coeff])]
        (.set desc P (long (+ 8 (* 8 k))) ^MemorySegment seg))
      (.set desc I 56 (int (if (= precision :c128) qsv-c128 qsv-c64)))
      (check (call qsv-batch-expect-z desc out))
      (mapv #(.getAtIndex out D (long %)) (range cnt)))))
